// pdk_force.hip -- PD bond passes with the velocity-Verlet update fused into
// the force epilogue.
//
// Mapping: one wavefront per 64-row slice of the sliced CSR, one lane per
// particle, a serial loop over the lane's bonds, no atomics -- the reference's
// SerialOpTag traversal (CabanaPD_Neighbor.hpp:31,134-151) laid out for
// 64-wide wavefronts: each slot row of 64 neighbor ids is one coalesced 256 B
// load and the broken-bond word of a slot row is one uniform 8 B load.
//
// Precision forms:
//   FP64  : the reference's expressions, operand order and rounding
//           (contraction off, correctly rounded sqrt/div).  Rows are summed
//           in the oracle's order, so f/u/v match the oracle bit for bit.
//   MIXED : bond geometry in FP32 from FP64 differences (no large-coordinate
//           cancellation: xi = x_j - x_i and eta = u_j - u_i are formed in
//           FP64 and rounded once), stretch s = (2 xi.eta + eta.eta) /
//           (|xi| (|xi+eta| + |xi|)), bond forces accumulated in FP64.  A
//           bond whose FP32 stretch lies inside a guard band around s0 is
//           re-tested with the FP64 expression, so break decisions equal the
//           FP64 decisions for the same state.
#include <cmath>
#include <cstdio>

#include "pdk_internal.hpp"

namespace pdk
{
namespace
{
// Values equal the ABI's model kinds.  The linearized models take the force
// model of their parent (LinearPMB: PMB, LinearLPS: LPS) and change only the
// bond geometry of the force and energy passes.
enum ModelT : int
{
    kPMB = 0,
    kLPS = 1,
    kLinPMB = 4,
    kLinLPS = 5
};
template <int M>
constexpr bool kIsLps = ( M == kLPS || M == kLinLPS );
template <int M>
constexpr bool kIsLin = ( M == kLinPMB || M == kLinLPS );
enum PrecT : int
{
    kF64 = 0,
    kMixed = 1
};

struct Bond64
{
    double xi, r, s, rx, ry, rz;
};

// getDistance (CabanaPD_Force.hpp:74-93), operation for operation.
__device__ __forceinline__ void getDistance64( double xi0, double xi1, double xi2, double ui0,
                                               double ui1, double ui2, double xj0, double xj1,
                                               double xj2, double uj0, double uj1, double uj2,
                                               Bond64& b )
{
#pragma clang fp contract( off )
    const double xi_x = xj0 - xi0;
    const double eta_u = uj0 - ui0;
    const double xi_y = xj1 - xi1;
    const double eta_v = uj1 - ui1;
    const double xi_z = xj2 - xi2;
    const double eta_w = uj2 - ui2;
    b.rx = xi_x + eta_u;
    b.ry = xi_y + eta_v;
    b.rz = xi_z + eta_w;
    b.r = sqrt( b.rx * b.rx + b.ry * b.ry + b.rz * b.rz );
    b.xi = sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
    b.s = ( b.r - b.xi ) / b.xi;
}

// BaseFractureModel CriticalStretchTag (CabanaPD_ForceModels.hpp:123-128).
__device__ __forceinline__ bool criticalStretch64( double r, double xi, double bcc )
{
#pragma clang fp contract( off )
    return r * r >= bcc * xi * xi;
}

// TemperatureDependent (CabanaPD_ForceModels.hpp:203-211): the bond's
// temp_avg; the stretch becomes s - alpha * temp_avg.
__device__ __forceinline__ double thermalAvg( const PDArgs& A, double Ti, double Tj )
{
#pragma clang fp contract( off )
    return 0.5 * ( Ti + Tj ) - A.temp0;
}
// ThermalFractureModel CriticalStretchTag coefficient (CabanaPD_ForceModels.hpp:246-255).
__device__ __forceinline__ double thermalBcc( const PDArgs& A, double temp_avg )
{
#pragma clang fp contract( off )
    return ( 1.0 + A.s0 + A.alpha * temp_avg ) * ( 1.0 + A.s0 + A.alpha * temp_avg );
}
// EPP ForceCoeffTag (force_models/CabanaPD_PMB.hpp:104-121): update the bond's
// plastic stretch (written only when it changes) and return s - s_p.
__device__ __forceinline__ double plasticStretch( double* slot, double s, double sY )
{
#pragma clang fp contract( off )
    double sp = *slot;
    if ( s >= sp + sY )
    {
        sp = s - sY;
        *slot = sp;
    }
    else if ( s <= sp - sY )
    {
        sp = s + sY;
        *slot = sp;
    }
    return s - sp;
}
// The same with the slot's stored plastic stretch already loaded (sp).
__device__ __forceinline__ double plasticStretchAt( double* slot, double sp, double s, double sY )
{
#pragma clang fp contract( off )
    if ( s >= sp + sY )
    {
        sp = s - sY;
        *slot = sp;
    }
    else if ( s <= sp - sY )
    {
        sp = s + sY;
        *slot = sp;
    }
    return s - sp;
}
// EPP EnergyTag stretch term (force_models/CabanaPD_PMB.hpp:123-146).
__device__ __forceinline__ double plasticEnergyTerm( double s, double sp, double sY )
{
#pragma clang fp contract( off )
    if ( s >= sp + sY )
        return sY * ( 2.0 * s - sY );
    if ( s <= sp - sY )
        return sY * ( -2.0 * s - sY );
    return s * s;
}

struct Bond32
{
    float s, rinv, rx, ry, rz, xi;
};

// Mixed-form bond geometry from the FP64 bond vector r = xi + (u_j - u_i)
// (the reference bond xi of the stencil entry, or x_j - x_i, plus the
// displacement difference, so no large coordinate cancels): the stretch
// numerator |r|^2 - |xi|^2 in FP64 as one FMA chain seeded with -|xi|^2 (the
// FP64 square of the reference bond), then 1/|r| and
//   s = (|r|^2 - |xi|^2) / (|xi| (|r| + |xi|))
// in FP32, with |r|^2 = |xi|^2 + numerator.  The numerator is exact to FP64
// rounding for any rotation of the bond (no xi.eta against eta.eta
// cancellation in FP32), so the FP32 stretch is within a few ulp of s.  The
// force passes accumulate the bond force along the FP64 r.
__device__ __forceinline__ void mixedBond( double r0, double r1, double r2, double xi2d, float xi, Bond32& b )
{
    const float num = static_cast<float>( fma( r0, r0, fma( r1, r1, fma( r2, r2, -xi2d ) ) ) );
    const float r2f = fmaf( xi, xi, num );
    b.rinv = __builtin_amdgcn_rsqf( r2f );
    b.s = num * __builtin_amdgcn_rcpf( fmaf( r2f, b.rinv, xi ) * xi );
    b.rx = static_cast<float>( r0 );
    b.ry = static_cast<float>( r1 );
    b.rz = static_cast<float>( r2 );
    b.xi = xi;
}

// Wave-uniform read through the scalar data cache (p must be the same in
// every lane).  Used for the stencil tables, which no kernel writes, and for
// the per-slot bond-mask words, which a wave reads before it writes them (a
// vector store) and which no other wave of the launch touches; the scalar
// cache starts every kernel dispatch invalidated.
template <class T>
__device__ __forceinline__ T scalarLoad( const T* p )
{
    return *( const __attribute__( ( address_space( 4 ) ) ) T* )( p );
}
__device__ __forceinline__ float4 scalarLoad4( const float4* p )
{
    const auto* f = ( const __attribute__( ( address_space( 4 ) ) ) float* )( p );
    return make_float4( f[0], f[1], f[2], f[3] );
}
// Raw buffer access for the row windows: an SGPR resource (uniform base and
// byte size), a VGPR byte offset (the lane) and an SGPR byte offset (the
// window start), so no address arithmetic runs on the vector ALUs.  Offsets
// at or past the size read 0 (hardware range check), which also covers
// windows that start before id 0 (their negative offsets wrap to values
// above the size for any specimen of fewer than 2^28 particles).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bufferOf( const void* base, uint32_t bytes )
{
    // base is wave-uniform; readfirstlane keeps the resource in SGPRs
    const uint64_t b = reinterpret_cast<uint64_t>( base );
    const uint32_t lo = __builtin_amdgcn_readfirstlane( static_cast<uint32_t>( b ) );
    const uint32_t hi = __builtin_amdgcn_readfirstlane( static_cast<uint32_t>( b >> 32 ) );
    void* ub = reinterpret_cast<void*>( ( static_cast<uint64_t>( hi ) << 32 ) | lo );
    // dword 3 of a raw (untyped) buffer resource on gfx9-family GPUs
    return __builtin_amdgcn_make_buffer_rsrc( ub, 0, static_cast<int>( __builtin_amdgcn_readfirstlane( bytes ) ),
                                              0x00020000 );
}
__device__ __forceinline__ double bufferLoad64( __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff )
{
    return __builtin_bit_cast( double, __builtin_amdgcn_raw_buffer_load_b64( r, voff, soff, 0 ) );
}
__device__ __forceinline__ uint8_t bufferLoad8( __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff )
{
    return __builtin_amdgcn_raw_buffer_load_b8( r, voff, soff, 0 );
}

__device__ __forceinline__ int4 scalarLoad4( const int4* p )
{
    const auto* f = ( const __attribute__( ( address_space( 4 ) ) ) int* )( p );
    return make_int4( f[0], f[1], f[2], f[3] );
}

// getLinearizedDistance (CabanaPD_Force.hpp:114-138), operation for
// operation: stretch linearized in eta, force along the reference bond.
// Stored so that the force passes' coeff * rx / r is coeff * xi_x / xi.
__device__ __forceinline__ void getLinearizedDistance64( double xi0, double xi1, double xi2, double ui0,
                                                         double ui1, double ui2, double xj0, double xj1,
                                                         double xj2, double uj0, double uj1, double uj2,
                                                         Bond64& b )
{
#pragma clang fp contract( off )
    const double xi_x = xj0 - xi0;
    const double eta_u = uj0 - ui0;
    const double xi_y = xj1 - xi1;
    const double eta_v = uj1 - ui1;
    const double xi_z = xj2 - xi2;
    const double eta_w = uj2 - ui2;
    b.xi = sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
    b.s = ( xi_x * eta_u + xi_y * eta_v + xi_z * eta_w ) / ( b.xi * b.xi );
    b.rx = xi_x;
    b.ry = xi_y;
    b.rz = xi_z;
    b.r = b.xi;
}

// Mixed form of the linearized stretch: s = xi.eta / |xi|^2 in FP32 from the
// FP64 differences; direction and 1/r are those of the reference bond.
__device__ __forceinline__ void linearized32( float gx, float gy, float gz, float xi, float xinv, double ui0,
                                              double ui1, double ui2, double uj0, double uj1, double uj2, Bond32& b )
{
    const float ex = static_cast<float>( uj0 - ui0 );
    const float ey = static_cast<float>( uj1 - ui1 );
    const float ez = static_cast<float>( uj2 - ui2 );
    b.rx = gx;
    b.ry = gy;
    b.rz = gz;
    b.xi = xi;
    b.rinv = xinv;
    b.s = ( gx * ex + gy * ey + gz * ez ) * ( xinv * xinv );
}

// Bond geometry of the force and energy passes for the model.
template <int MODEL>
__device__ __forceinline__ void bond64( double xi0, double xi1, double xi2, double ui0, double ui1, double ui2,
                                        double xj0, double xj1, double xj2, double uj0, double uj1, double uj2,
                                        Bond64& b )
{
    if constexpr ( kIsLin<MODEL> )
        getLinearizedDistance64( xi0, xi1, xi2, ui0, ui1, ui2, xj0, xj1, xj2, uj0, uj1, uj2, b );
    else
        getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, xj0, xj1, xj2, uj0, uj1, uj2, b );
}
// Linearized stretch of the mixed form, bond vector gathered (x_j - x_i)
// or from the stencil table (g = (xi_x, xi_y, xi_z, |xi|), xinv = 1/|xi|).
__device__ __forceinline__ void linearized32g( double xi0, double xi1, double xi2, double ui0, double ui1,
                                               double ui2, double xj0, double xj1, double xj2, double uj0,
                                               double uj1, double uj2, Bond32& b )
{
    const float gx = static_cast<float>( xj0 - xi0 );
    const float gy = static_cast<float>( xj1 - xi1 );
    const float gz = static_cast<float>( xj2 - xi2 );
    const float g2 = gx * gx + gy * gy + gz * gz;
    const float xinv = __builtin_amdgcn_rsqf( g2 );
    linearized32( gx, gy, gz, g2 * xinv, xinv, ui0, ui1, ui2, uj0, uj1, uj2, b );
}

// Uniform per-wave access to the spec of this wave's slice.
struct WaveCtx
{
    const PDArgs* A;
    int64_t slice;
    int32_t epi;
    int32_t cur;
};

__device__ __forceinline__ bool waveSetup( const PDArgs* __restrict__ specs,
                                           const int32_t* __restrict__ tile_spec,
                                           const int64_t* __restrict__ tile_slice, int32_t epi,
                                           int32_t parity, int64_t launch, WaveCtx& w )
{
    // uniform per wavefront; readfirstlane lets the compiler keep it (and the
    // addresses derived from it) in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int64_t t = xcdRemap( blockIdx.x, gridDim.x );
    if ( tile_spec == nullptr )
    {
        w.A = specs;
        w.slice = t * kSlicesPerBlock + wave;
        w.epi = epi;
        w.cur = parity;
    }
    else
    {
        // the tile tables are written by the host before the launch
        w.A = specs + scalarLoad( tile_spec + t );
        w.slice = scalarLoad( tile_slice + t ) + wave;
        // launch < 0: step -1 - launch of a replayed graph, whose first step's
        // index is the device counter (written by an earlier kernel of the
        // stream; the scalar cache starts every dispatch invalidated)
        const int64_t L = launch >= 0 ? launch : scalarLoad( scalarLoad( &w.A->launch_ctr ) ) + ( -1 - launch );
        w.cur = static_cast<int32_t>( L & 1 ) ^ w.A->parity_offset;
        w.epi = ( L + 1 < w.A->end_launch ) ? kEpiFinalInitial : kEpiFinal;
    }
    return w.slice < w.A->nslices;
}

// Force BC and body term (CabanaPD_Boundary.hpp:94-157, CabanaPD_BodyTerm.hpp:46-63),
// then the velocity-Verlet halves (CabanaPD_Integrate.hpp:100-109, 128-134).
__device__ __forceinline__ void epilogue( const PDArgs& A, int64_t p, int32_t epi, int32_t cur,
                                          double fx, double fy, double fz )
{
#pragma clang fp contract( off )
    const int64_t n = A.n;
    if ( epi != kEpiStoreOnly )
    {
        if ( A.body )
        {
            fx = fx + A.body[p];
            fy = fy + A.body[n + p];
            fz = fz + A.body[2 * n + p];
        }
        if ( A.bc_kind >= 0 && A.bcflag[p] )
        {
            if ( A.bc_kind == PDK_BC_FORCE_VALUE )
            {
                fx = A.bc_value;
                fy = A.bc_value;
                fz = A.bc_value;
            }
            else
            {
                fx = fx + A.bc_value;
                fy = fy + A.bc_value;
                fz = fz + A.bc_value;
            }
        }
    }
    A.f[p] = fx;
    A.f[n + p] = fy;
    A.f[2 * n + p] = fz;
    if ( epi == kEpiStoreOnly )
        return;
    const double half_dt_m = A.half_dt / A.rho[p];
    double vx = A.v[p], vy = A.v[n + p], vz = A.v[2 * n + p];
    vx = vx + half_dt_m * fx;
    vy = vy + half_dt_m * fy;
    vz = vz + half_dt_m * fz;
    if ( epi == kEpiFinalInitial )
    {
        vx = vx + half_dt_m * fx;
        vy = vy + half_dt_m * fy;
        vz = vz + half_dt_m * fz;
        const double* u = A.U[cur];
        double* un = A.U[cur ^ 1];
        un[p] = u[p] + A.dt * vx;
        un[n + p] = u[n + p] + A.dt * vy;
        un[2 * n + p] = u[2 * n + p] + A.dt * vz;
    }
    A.v[p] = vx;
    A.v[n + p] = vy;
    A.v[2 * n + p] = vz;
}

// ---------------------------------------------------------------------------
// Force pass: Force<PMB|LPS, NoFracture|Fracture>::computeForceFull
// (force/CabanaPD_PMB.hpp:89-125,222-287; force/CabanaPD_LPS.hpp:143-181,380-447)
// For LPS the state pass has stored A = theta_coeff*theta/m and B = s_coeff/m
// per particle (bitwise the per-bond subexpressions of the reference's
// ForceCoeffTag, force_models/CabanaPD_LPS.hpp:131-133).
// ---------------------------------------------------------------------------
// SLOTS selects the stencil-slot form (neighbor ids from the slot offsets,
// existence bits per slot row) over the compact form (neighbor id stream).
// MM: the mixed form of a multi-material specimen (pair tables), an instance
// of its own so that the one-material mixed instances carry none of its
// per-bond table lookups (the FP64 form reads the tables when present).
// ---------------------------------------------------------------------------
constexpr int kBatch = 16; // bond slots whose ids / bit words are fetched ahead

template <int MODEL, bool FRAC, int PREC, int UNR, bool SLOTS, bool MM = false>
__global__ __launch_bounds__( kBlock ) __attribute__( ( amdgpu_waves_per_eu( ( PREC == kMixed && !SLOTS && UNR <= 2 ) ? 4 : 1, 8 ) ) ) void kForce( const PDArgs* __restrict__ specs,
                                                    const int32_t* __restrict__ tile_spec,
                                                    const int64_t* __restrict__ tile_slice,
                                                    int32_t epi, int32_t parity, int64_t launch,
                                                    int32_t accumulate )
{
#pragma clang fp contract( off )
    WaveCtx w;
    const bool active = waveSetup( specs, tile_spec, tile_slice, epi, parity, launch, w );
    // All waves of a block belong to one spec: stage its stencil table (and
    // slot offsets) in LDS before any wave leaves.
    __shared__ float4 stab[( PREC == kMixed ) ? 256 : 1];
    __shared__ double sxi2[( PREC == kMixed ) ? 256 : 1]; // FP64 |xi|^2 of the stencil entries
    __shared__ int32_t soffs[SLOTS ? 256 + kBatch : 1];
    __shared__ uint32_t rsm[( PREC == kMixed ) ? 9 : 1]; // row-start bits of the slots (stencil-slot form)
    // compact form with stencil ids: the FP64 record of every stencil entry
    // (kXiX, kXi2, kXiY, kXiZ), so that the mixed form reads xi instead of x_j
    __shared__ __align__( 16 ) double sxv[( PREC == kMixed && !SLOTS ) ? 4 * 256 : 1];
    const bool useSten = ( PREC == kMixed ) && ( SLOTS || w.A->sid != nullptr );
    if constexpr ( PREC == kMixed )
    {
        if ( threadIdx.x < 9 )
            rsm[threadIdx.x] = ( SLOTS && w.A->rowstart ) ? w.A->rowstart[threadIdx.x] : 0xffffffffu;
    }
    if constexpr ( SLOTS )
    {
        soffs[threadIdx.x] = w.A->soff[threadIdx.x];
        if ( threadIdx.x < kBatch )
            soffs[256 + threadIdx.x] = 0;
    }
    if constexpr ( PREC == kMixed )
    {
        if ( useSten )
        {
            stab[threadIdx.x] = w.A->stencil[threadIdx.x];
            sxi2[threadIdx.x] = w.A->stencil_xi2d[threadIdx.x];
            if constexpr ( !SLOTS )
            {
#pragma unroll
                for ( int k = 0; k < 4; k++ )
                    sxv[4 * threadIdx.x + k] = w.A->stencil_xiv[4 * threadIdx.x + k];
            }
        }
    }
    if constexpr ( SLOTS || PREC == kMixed )
        __syncthreads();
    if ( !active )
        return;
    const PDArgs& A = *w.A;
    const double* __restrict__ sxiv = scalarLoad( &A.stencil_xiv );
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t n = A.n;
    const int64_t p = w.slice * kSlice + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : w.slice * kSlice;
    const double* __restrict__ x = A.x;
    const double* __restrict__ u = A.U[w.cur];
    const double xi0 = x[pc], xi1 = x[n + pc], xi2 = x[2 * n + pc];
    const double ui0 = u[pc], ui1 = u[n + pc], ui2 = u[2 * n + pc];
    const uint8_t* __restrict__ sidp = nullptr;
    const int32_t mycnt = valid ? A.cnt[p] : 0;
    const int64_t base = A.slice_ptr[w.slice];
    const int32_t len = static_cast<int32_t>( A.slice_ptr[w.slice + 1] - base );
    const int32_t* __restrict__ colp = A.col + base * kSlice + lane;
    if ( useSten && !SLOTS )
        sidp = A.sid + base * kSlice + lane;
    uint64_t* __restrict__ bits = A.intact + base;
    const double* __restrict__ vol = A.vol;
    const bool nf_i = FRAC ? ( A.nofail[pc] != 0 ) : false;
    const double c = A.c;
    const double bcc = A.bcc;
    // multi-material: model of the material pair; null with one material
    const uint8_t* __restrict__ mt = ( PREC == kF64 || MM ) ? A.mtype : nullptr;
    const int ti4 = mt ? 4 * static_cast<int>( mt[pc] ) : 0;
    // PMB family extensions: temperature, plastic stretch per bond slot
    const double* __restrict__ tmp = kIsLps<MODEL> ? nullptr : A.temp;
    double* __restrict__ spl = kIsLps<MODEL> ? nullptr : A.splas;
    const double Ti = tmp ? tmp[pc] : 0.0;
    double Ai = 0.0, Bi = 0.0;
    if constexpr ( kIsLps<MODEL> )
    {
        Ai = A.lpsA[pc]; // theta_coeff * theta_i / m_i
        Bi = A.lpsB[pc]; // s_coeff / m_i
    }
    const double* __restrict__ Aarr = A.lpsA;
    const double* __restrict__ Barr = A.lpsB;
    const int infl1 = A.influence;

    const bool uvol = A.vol_uniform != 0;
    const double volu = A.vol_value;
    const bool any_nofail = A.any_nofail != 0;

    double fx = 0.0, fy = 0.0, fz = 0.0;
    if ( accumulate && valid )
    {
        fx = A.f[p];
        fy = A.f[n + p];
        fz = A.f[2 * n + p];
    }
    // Mixed form: bond forces are summed in FP32 within a lattice row of the
    // stencil (at most kRowMaxSlots bonds) and the row sums in FP64; without a
    // row table every bond is its own row.
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    // non-linear mixed form: the bond forces fr r (FP32 fr, FP64 bond vector r)
    // summed in FP64, as in the row-window pass
    constexpr bool XV = ( PREC == kMixed ) && !kIsLin<MODEL>;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    // Uniform volume (the specimens the row-window pass takes without its
    // general path): the per-bond factor c vol (PMB) or vol (LPS) multiplies
    // each row sum in FP64 instead of every bond.
    // (PMB with several materials: c varies per pair, so it stays in each bond)
    const bool fold = uvol && !( FRAC && any_nofail ) && !( mt && !kIsLps<MODEL> );
    const double rscale = fold ? A.row_scale : 1.0;
    const float c32 = static_cast<float>( c );
    const float bandK = A.bandK;
    // The neighbor ids and broken-bond words of KB slots are fetched one
    // batch ahead (the index stream comes from HBM; keeping a batch in
    // flight hides its latency).  Within a batch, slots are processed UNR at
    // a time: the gathers of a group are issued together, then its bonds are
    // evaluated in slot order (the summation order is unchanged).
    // (compact mixed form: batches of 4 slots (or one group), which leave
    // the registers for 4 waves per SIMD: 0.997 -> 0.788 ms on a 1.44M
    // particle plate; the FP64 form keeps 16, it runs slower with 4)
    constexpr int KB = ( PREC == kMixed && !SLOTS ) ? ( UNR > 4 ? UNR : 4 ) : kBatch;
    static_assert( KB % UNR == 0, "batch must hold whole groups" );
    int32_t jb[SLOTS ? 1 : KB], jn[SLOTS ? 1 : KB];
    uint32_t sb[KB / 4], sn[KB / 4]; // stencil ids, four per word
    uint32_t wlo = 0, whi = 0, nlo = 0, nhi = 0;
    uint32_t elo = 0, ehi = 0, mlo = 0, mhi = 0; // existence words (stencil-slot form)
    const int32_t self = static_cast<int32_t>( pc );
    const uint64_t* __restrict__ exist = SLOTS ? A.exist + base : nullptr;
    auto fetch = [&]( int32_t qb, int32_t( &dst )[SLOTS ? 1 : KB], uint32_t( &sdst )[KB / 4], uint32_t& lo,
                      uint32_t& hi, uint32_t& xlo, uint32_t& xhi ) {
        if constexpr ( SLOTS )
        {
            // lane k holds the existence word of slot qb + k
            const uint64_t xv = ( lane < KB && qb + lane < len ) ? exist[qb + lane] : 0ull;
            xlo = static_cast<uint32_t>( xv );
            xhi = static_cast<uint32_t>( xv >> 32 );
        }
        else
        {
#pragma unroll
            for ( int k = 0; k < KB; k++ )
                dst[k] = ( qb + k < len ) ? colp[( qb + k ) * kSlice] : self;
        }
        if constexpr ( PREC == kMixed && !SLOTS )
        {
            if ( useSten )
            {
#pragma unroll
                for ( int k4 = 0; k4 < KB / 4; k4++ )
                {
                    uint32_t pk = 0;
#pragma unroll
                    for ( int e = 0; e < 4; e++ )
                    {
                        const int32_t q = qb + 4 * k4 + e;
                        const uint32_t v = ( q < len ) ? sidp[q * kSlice] : 0u;
                        pk |= v << ( 8 * e );
                    }
                    sdst[k4] = pk;
                }
            }
        }
        if constexpr ( FRAC )
        {
            // lane k holds the word of slot qb + k
            const uint64_t wv = ( lane < KB && qb + lane < len ) ? bits[qb + lane] : 0ull;
            lo = static_cast<uint32_t>( wv );
            hi = static_cast<uint32_t>( wv >> 32 );
        }
    };
    fetch( 0, jb, sb, wlo, whi, elo, ehi );
    for ( int32_t qb = 0; qb < len; qb += KB )
    {
        if ( qb + KB < len )
            fetch( qb + KB, jn, sn, nlo, nhi, mlo, mhi );
#pragma unroll
        for ( int g = 0; g < KB; g += UNR )
        {
        const int32_t q0 = qb + g;
        if ( q0 >= len )
            break;
        int64_t jj[UNR];
        uint64_t ww[UNR];
        bool ex[UNR];
#pragma unroll
        for ( int k = 0; k < UNR; k++ )
        {
            if constexpr ( SLOTS )
            {
                const uint32_t l32 = __builtin_amdgcn_readlane( elo, g + k );
                const uint32_t h32 = __builtin_amdgcn_readlane( ehi, g + k );
                const uint64_t e = ( static_cast<uint64_t>( h32 ) << 32 ) | l32;
                ex[k] = ( e >> lane ) & 1ull;
                jj[k] = ex[k] ? p + soffs[q0 + k] : pc;
            }
            else
            {
                jj[k] = jb[g + k];
                ex[k] = q0 + k < mycnt;
            }
            ww[k] = 0;
            if constexpr ( FRAC )
            {
                const uint32_t l32 = __builtin_amdgcn_readlane( wlo, g + k );
                const uint32_t h32 = __builtin_amdgcn_readlane( whi, g + k );
                ww[k] = ( static_cast<uint64_t>( h32 ) << 32 ) | l32;
            }
        }
        double xj[UNR][3], uj[UNR][3], vj[UNR], aj[UNR], bj[UNR], tj[UNR];
        bool nfj[UNR];
        int mj[UNR];
#pragma unroll
        for ( int k = 0; k < UNR; k++ )
        {
            mj[k] = 0;
            tj[k] = 0.0;
            if ( ex[k] )
            {
                const int64_t j = jj[k];
                if ( mt )
                    mj[k] = mt[j];
                if ( tmp )
                    tj[k] = tmp[j];
                // x_j: the FP64 form and the layouts without a stencil; the
                // mixed forms on a lattice take xi from the stencil (the
                // guard band's FP64 retest loads x_j itself)
                if ( PREC == kF64 || !useSten )
                {
                    xj[k][0] = x[j];
                    xj[k][1] = x[n + j];
                    xj[k][2] = x[2 * n + j];
                }
                uj[k][0] = u[j];
                uj[k][1] = u[n + j];
                uj[k][2] = u[2 * n + j];
                vj[k] = uvol ? volu : vol[j];
                nfj[k] = ( FRAC && !kIsLps<MODEL> && any_nofail ) ? ( A.nofail[j] != 0 ) : false;
                if constexpr ( kIsLps<MODEL> )
                {
                    aj[k] = Aarr[j];
                    bj[k] = Barr[j];
                }
            }
        }
#pragma unroll
        for ( int k = 0; k < UNR; k++ )
        {
            const int32_t q = q0 + k;
            bool brk = false;
            if constexpr ( PREC == kMixed && !XV )
            {
                if ( ( rsm[q >> 5] >> ( q & 31 ) ) & 1u )
                {
                    fx = fma( rscale, static_cast<double>( px ), fx );
                    fy = fma( rscale, static_cast<double>( py ), fy );
                    fz = fma( rscale, static_cast<double>( pz ), fz );
                    px = py = pz = 0.0f;
                }
            }
            if ( ex[k] )
            {
                bool intact = true;
                if constexpr ( FRAC )
                    intact = ( ww[k] >> lane ) & 1ull;
                const bool nf_j = kIsLps<MODEL> ? nf_i : nfj[k];
                if constexpr ( PREC == kF64 )
                {
                    Bond64 b;
                    bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, xj[k][0], xj[k][1], xj[k][2], uj[k][0], uj[k][1],
                                   uj[k][2], b );
                    double temp_avg = 0.0;
                    if ( tmp )
                    {
                        temp_avg = thermalAvg( A, Ti, tj[k] );
                        b.s = b.s - ( A.alpha * temp_avg );
                    }
                    if constexpr ( FRAC )
                    {
                        const double bcc_ij = tmp ? thermalBcc( A, temp_avg ) : ( mt ? A.mat_bcc[ti4 + mj[k]] : bcc );
                        if ( criticalStretch64( b.r, b.xi, bcc_ij ) && !nf_i && !nf_j )
                            brk = true;
                    }
                    if ( !brk && intact )
                    {
                        double coeff;
                        if constexpr ( !kIsLps<MODEL> )
                        {
                            const double sd =
                                spl ? plasticStretch( spl + ( base + q ) * kSlice + lane, b.s, A.sY ) : b.s;
                            coeff = ( mt ? A.mat_c[ti4 + mj[k]] : c ) * sd * vj[k];
                        }
                        else
                        {
                            const double infl = infl1 ? 1.0 / b.xi : 1.0;
                            coeff = ( Ai + aj[k] + b.s * ( Bi + bj[k] ) ) * infl * b.xi * vj[k];
                        }
                        fx += coeff * b.rx / b.r;
                        fy += coeff * b.ry / b.r;
                        fz += coeff * b.rz / b.r;
                    }
                }
                else
                {
                    Bond32 b;
                    double r0 = 0.0, r1 = 0.0, r2 = 0.0; // XV: the FP64 bond vector
                    float wl = 1.0f; // LPS: infl |xi|
                    if constexpr ( kIsLin<MODEL> )
                    {
                        if ( useSten )
                        {
                            const uint32_t id = SLOTS ? static_cast<uint32_t>( q )
                                                      : ( sb[( g + k ) / 4] >> ( 8 * ( ( g + k ) % 4 ) ) ) & 0xffu;
                            const float4 g4 = stab[id];
                            linearized32( g4.x, g4.y, g4.z, g4.w, __builtin_amdgcn_rcpf( g4.w ), ui0, ui1, ui2,
                                          uj[k][0], uj[k][1], uj[k][2], b );
                        }
                        else
                            linearized32g( xi0, xi1, xi2, ui0, ui1, ui2, xj[k][0], xj[k][1], xj[k][2], uj[k][0],
                                           uj[k][1], uj[k][2], b );
                    }
                    else
                    {
                        // r = xi + (u_j - u_i): xi the slot's FP64 stencil offset
                        // (as in the row-window pass) or x_j - x_i
                        double xi2d;
                        float xil;
                        double g0, g1, g2;
                        if ( useSten )
                        {
                            const uint32_t id = SLOTS ? static_cast<uint32_t>( q )
                                                      : ( sb[( g + k ) / 4] >> ( 8 * ( ( g + k ) % 4 ) ) ) & 0xffu;
                            xi2d = sxi2[id];
                            xil = stab[id].w;
                            if constexpr ( SLOTS )
                            {
                                // slot q is wave-uniform: scalar loads of its FP64 record
                                const double* xv = sxiv + 4 * q;
                                g0 = scalarLoad( xv + kXiX );
                                g1 = scalarLoad( xv + kXiY );
                                g2 = scalarLoad( xv + kXiZ );
                            }
                            else
                            {
                                const double* xv = sxv + 4 * id;
                                g0 = xv[kXiX];
                                g1 = xv[kXiY];
                                g2 = xv[kXiZ];
                                xi2d = xv[kXi2];
                            }
                        }
                        else
                        {
                            g0 = xj[k][0] - xi0;
                            g1 = xj[k][1] - xi1;
                            g2 = xj[k][2] - xi2;
                            xi2d = fma( g0, g0, fma( g1, g1, g2 * g2 ) );
                            const float x2f = static_cast<float>( xi2d );
                            xil = x2f * __builtin_amdgcn_rsqf( x2f );
                        }
                        // (the y and z terms add the displacement of j last,
                        // so the row-window pass can form xi - u_i once per
                        // lattice row)
                        r0 = ( uj[k][0] - ui0 ) + g0;
                        r1 = uj[k][1] + ( g1 - ui1 );
                        r2 = uj[k][2] + ( g2 - ui2 );
                        mixedBond( r0, r1, r2, xi2d, xil, b );
                        wl = infl1 ? 1.0f : xil;
                    }
                    // TemperatureDependent: the thermal strain alpha * temp_avg
                    // (FP64, rounded once) leaves the stretch; the guard band
                    // then scales with the bond's terms, which bound the
                    // rounding of the mechanical stretch, the thermal strain
                    // and their difference.
                    double temp_avg = 0.0;
                    float ath = 0.0f;
                    const float s_mech = b.s;
                    if ( tmp )
                    {
                        temp_avg = thermalAvg( A, Ti, tj[k] );
                        ath = static_cast<float>( A.alpha * temp_avg );
                        b.s = b.s - ath;
                    }
                    if constexpr ( FRAC )
                    {
                        if ( !nf_i && !nf_j )
                        {
                            // the pair's critical stretch and band (several materials)
                            const float s0p = mt ? A.mat_s0f[ti4 + mj[k]] : A.s0f;
                            const float bandp = mt ? A.mat_bandK[ti4 + mj[k]] : bandK;
                            const float band =
                                tmp ? fmaf( A.bandf, fabsf( s_mech ) + s0p + fabsf( ath ), bandp ) : bandp;
                            if ( fabsf( b.s - s0p ) <= band )
                            {
                                // guard band: the reference's FP64 test on this bond
                                Bond64 e;
                                const int64_t j = jj[k];
                                const bool xl = PREC == kF64 || !useSten;
                                getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, xl ? xj[k][0] : x[j],
                                               xl ? xj[k][1] : x[n + j], xl ? xj[k][2] : x[2 * n + j], uj[k][0],
                                               uj[k][1], uj[k][2], e );
                                brk = criticalStretch64( e.r, e.xi,
                                                         tmp ? thermalBcc( A, temp_avg )
                                                             : ( mt ? A.mat_bcc[ti4 + mj[k]] : bcc ) );
                            }
                            else
                                brk = b.s > s0p;
                        }
                    }
                    if ( !brk && intact )
                    {
                        float fr;
                        if constexpr ( !kIsLps<MODEL> )
                        {
                            // ElasticPerfectlyPlastic: the plastic stretch stays FP64
                            const float sd =
                                spl ? static_cast<float>( plasticStretch( spl + ( base + q ) * kSlice + lane,
                                                                          static_cast<double>( b.s ), A.sY ) )
                                    : b.s;
                            const float cp = mt ? static_cast<float>( A.mat_c[ti4 + mj[k]] ) : c32;
                            fr = ( fold ? sd : ( cp * static_cast<float>( vj[k] ) ) * sd ) * b.rinv;
                        }
                        else
                        {
                            const float coeff =
                                static_cast<float>( Ai + aj[k] ) + b.s * static_cast<float>( Bi + bj[k] );
                            fr = ( fold ? coeff * wl : ( coeff * wl ) * static_cast<float>( vj[k] ) ) * b.rinv;
                        }
                        if constexpr ( XV )
                        {
                            const double frd = static_cast<double>( fr );
                            sx = fma( frd, r0, sx );
                            sy = fma( frd, r1, sy );
                            sz = fma( frd, r2, sz );
                        }
                        else
                        {
                            px = fmaf( fr, b.rx, px );
                            py = fmaf( fr, b.ry, py );
                            pz = fmaf( fr, b.rz, pz );
                        }
                    }
                }
            }
            if constexpr ( FRAC )
            {
                if ( q < len )
                {
                    const uint64_t m = __ballot( brk );
                    if ( ( m & ww[k] ) && lane == 0 )
                        bits[q] = ww[k] & ~m;
                }
            }
        }
        } // group
        if constexpr ( !SLOTS )
        {
#pragma unroll
            for ( int k = 0; k < KB; k++ )
                jb[k] = jn[k];
        }
#pragma unroll
        for ( int k4 = 0; k4 < KB / 4; k4++ )
            sb[k4] = sn[k4];
        wlo = nlo;
        whi = nhi;
        elo = mlo;
        ehi = mhi;
    }
    if constexpr ( PREC == kMixed && !XV )
    {
        fx = fma( rscale, static_cast<double>( px ), fx );
        fy = fma( rscale, static_cast<double>( py ), fy );
        fz = fma( rscale, static_cast<double>( pz ), fz );
    }
    if constexpr ( XV )
    {
        fx = fma( rscale, sx, fx );
        fy = fma( rscale, sy, fy );
        fz = fma( rscale, sz, fz );
    }
    if ( !valid )
        return;
    epilogue( A, p, w.epi, w.cur, fx, fy, fz );
}

// ---------------------------------------------------------------------------
// The same force pass with the neighbor data staged in LDS (stencil-slot form
// on lattice rows).  For a lattice row r = (dy, dz) of the stencil, slots
// q0 .. q0+c-1 of the 64 lanes of a slice have their neighbors at the
// contiguous ids g .. g+63+c-1 (g = first id of the slice + the row's offset):
// the wave loads that window once with coalesced loads, stores it in its LDS
// buffer, and slot q0+t of lane L reads window[L + t].  The window of the next
// row is loaded while the current row is evaluated.  Per slot, one
// wave-uniform 64-bit word (scalar load) gates the lanes: the intact word
// under fracture (breaks and pre-notches only clear bits, so it is a subset
// of the existence word) or the existence word.  Evaluation order and
// arithmetic are kForce's, so results are identical.
// ---------------------------------------------------------------------------
// Waves per SIMD the register allocator must leave room for (the attribute
// below): 6 for the PMB passes (80 VGPRs, one spill outside the loops;
// measured 6 % faster on the fracture plate than the unconstrained
// allocation at 5 waves); LPS (models 1 and 5) stays unconstrained, since
// forcing it spills inside the slot loop (mixed LPS at 6 waves: 8 VGPRs
// spilled, 0.349 -> 0.354 ms on a 128^3 cube).
template <int MODEL, bool FRAC, int PREC, bool GEN>
__global__ __launch_bounds__( kBlock ) __attribute__( ( amdgpu_waves_per_eu( ( MODEL == 1 || MODEL == 5 ) ? 1 : 6, 8 ) ) ) void
kForceRows( const PDArgs* __restrict__ specs,
                                                        const int32_t* __restrict__ tile_spec,
                                                        const int64_t* __restrict__ tile_slice,
                                                        int32_t epi, int32_t parity, int64_t launch,
                                                        int32_t accumulate )
{
#pragma clang fp contract( off )
    constexpr int kW = kSlice + kRowMaxSlots - 1;  // window width
    constexpr int NX = ( PREC == kF64 ) ? 3 : 0;    // staged x components
    // mixed form (non-linear stretch): the bond vector r = xi + (u_j - u_i)
    // in FP64 from the slot's FP64 stencil offset xi; the windows hold u only
    constexpr bool XV = ( PREC == kMixed ) && !kIsLin<MODEL>;
    constexpr int NAB = kIsLps<MODEL> ? 2 : 0;  // staged lpsA, lpsB
    constexpr int KV = 3 + NX + NAB;                // staged vol (GEN)
    constexpr int KNF = KV + 1;                     // staged no-fail flag (GEN)
    constexpr int NC = KV + ( GEN ? 2 : 0 );
    // channel stride: PMB mixed rows pad each channel to 128 values, so the
    // second window half is stored by every lane (no per-row store masks);
    // values past the row's window are never read
    constexpr bool PAD = ( PREC == kMixed ) && !kIsLps<MODEL> && !GEN;
    constexpr int kWS = PAD ? 2 * kSlice : kW;
    __shared__ double win[kSlicesPerBlock][2][NC][kWS];
    WaveCtx w;
    if ( !waveSetup( specs, tile_spec, tile_slice, epi, parity, launch, w ) )
        return; // waves are independent: no block-level synchronisation
    const PDArgs& A = *w.A;
    const int wv = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t n = scalarLoad( &A.n );
    const int64_t p0 = w.slice * kSlice;
    const int64_t p = p0 + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : p0;
    // Base pointers, the row table and the slice table through scalar loads
    // (the argument block and the tables are written by the host only), so
    // window addresses are an SGPR base plus the lane offset.
    const double* __restrict__ x = scalarLoad( &A.x );
    const double* __restrict__ u = scalarLoad( &A.U[w.cur] );
    const int4* __restrict__ rows = scalarLoad( &A.rows );
    // particle i: reference position (FP64 form) and displacement
    double xi0 = 0.0, xi1 = 0.0, xi2 = 0.0;
    if constexpr ( PREC == kF64 )
    {
        xi0 = x[pc];
        xi1 = x[n + pc];
        xi2 = x[2 * n + pc];
    }
    const double ui0 = u[pc], ui1 = u[n + pc], ui2 = u[2 * n + pc];
    // Per-slot lane masks of the slice: the intact words (a subset of the
    // existence words, so they alone gate the fracture forms) or, without
    // fracture, the existence words.  Lanes past n hold no bonds.
    // (the argument block and the slice table are written by the host only)
    const int64_t row0 = scalarLoad( scalarLoad( &A.slice_ptr ) + w.slice );
    uint64_t* __restrict__ bits = scalarLoad( &A.intact ) + row0;
    const uint64_t* __restrict__ mwords = FRAC ? bits : scalarLoad( &A.exist ) + row0;
    const float4* __restrict__ sten = scalarLoad( &A.stencil );
    const float4* __restrict__ saux = scalarLoad( &A.stencil_aux );
    const double* __restrict__ sxiv = scalarLoad( &A.stencil_xiv );
    const bool nf_i = FRAC ? ( A.nofail[pc] != 0 ) : false;
    // Loop-invariant scalars held in registers (the args block lives in
    // global memory, which the bit-word stores could alias).
    const double c = A.c;
    const double bcc = A.bcc;
    const float c32 = static_cast<float>( A.c );
    const double volu = A.vol_value;
    const float s0f = A.s0f;
    const float bandK = A.bandK;
    // lower edge of the decision path: every stretch within bandK of s0 or
    // above it lies at or above s0 - 2 bandK
    const float s0lo = s0f - 2.0f * bandK;
    // uniform volume: c vol (PMB) or vol (LPS) scales each FP32 row sum
    const double rscale = GEN ? 1.0 : A.row_scale;
    const int infl1 = A.influence;
    // lanes whose particle may break bonds (every lane of the wave is active here)
    const uint64_t nfi_ok = FRAC ? __ballot( !nf_i ) : 0ull;
    double Ai = 0.0, Bi = 0.0;
    if constexpr ( kIsLps<MODEL> )
    {
        Ai = A.lpsA[pc]; // theta_coeff * theta_i / m_i
        Bi = A.lpsB[pc]; // s_coeff / m_i
    }

    double fx = 0.0, fy = 0.0, fz = 0.0;
    if ( accumulate && valid )
    {
        fx = A.f[p];
        fy = A.f[n + p];
        fz = A.f[2 * n + p];
    }
    // XV: the bond forces fr r (FP32 fr, FP64 bond vector r) summed in FP64
    double sx = 0.0, sy = 0.0, sz = 0.0;

    const double* src[KV];
    src[0] = u;
    src[1] = u + n;
    src[2] = u + 2 * n;
    if constexpr ( NX )
    {
        src[3] = x;
        src[4] = x + n;
        src[5] = x + 2 * n;
    }
    if constexpr ( NAB )
    {
        src[3 + NX] = scalarLoad( &A.lpsA );
        src[4 + NX] = scalarLoad( &A.lpsB );
    }
    const double* __restrict__ volp = scalarLoad( &A.vol );
    const uint8_t* __restrict__ nfp = scalarLoad( &A.nofail );
    double v0[NC], v1[NC];
#pragma unroll
    for ( int k = 0; k < NC; k++ )
        v1[k] = 0.0;
    // Row tables exist only for n < 2^28 (byte sizes and offsets fit 32 bits).
    const int32_t p032 = static_cast<int32_t>( p0 );
    const uint32_t nbytes = static_cast<uint32_t>( n ) * 8u;
    __amdgpu_buffer_rsrc_t rs[KV + ( GEN ? 1 : 0 )];
#pragma unroll
    for ( int k = 0; k < KV; k++ )
        rs[k] = bufferOf( src[k], nbytes );
    const __amdgpu_buffer_rsrc_t rnf = bufferOf( nfp, GEN ? static_cast<uint32_t>( n ) : 0u );
    if constexpr ( GEN )
        rs[KV] = bufferOf( volp, nbytes );
    const uint32_t lane8 = static_cast<uint32_t>( lane ) * 8u;
    // 64 window values starting at id g (uniform) into dst.
    auto loadHalf = [&]( double* dst, int32_t g ) {
        // g >= 0: lane offset + uniform start; g < 0: per-lane offsets, the
        // negative ones wrap out of range and read 0
        const uint32_t vo = g >= 0 ? lane8 : static_cast<uint32_t>( g + lane ) * 8u;
        const uint32_t so = g >= 0 ? static_cast<uint32_t>( g ) * 8u : 0u;
#pragma unroll
        for ( int k = 0; k < KV; k++ )
            dst[k] = bufferLoad64( rs[k], vo, so );
        if constexpr ( GEN )
        {
            dst[KV] = bufferLoad64( rs[KV], vo, so );
            dst[KNF] = ( FRAC && bufferLoad8( rnf, vo >> 3, so >> 3 ) != 0 ) ? 1.0 : 0.0;
        }
    };
    // Window of row r: ids g .. g + 63 (every lane) and g + 64 .. g + 64 + c - 2
    // (lanes < c - 1).
    auto loadRow = [&]( int r ) {
        const int4 R = scalarLoad4( rows + r );
        const int32_t g = p032 + R.z; // uniform
        loadHalf( v0, g );
        if ( lane < R.y - 1 )
            loadHalf( v1, g + kSlice );
    };

    const int nrows = scalarLoad( &A.nrows );
    loadRow( 0 );
    for ( int r = 0; r < nrows; r++ )
    {
        const int4 R = scalarLoad4( rows + r );
        double* __restrict__ wb = &win[wv][r & 1][0][0];
#pragma unroll
        for ( int k = 0; k < NC; k++ )
        {
            wb[k * kWS + lane] = v0[k];
            if ( PAD || lane < R.y - 1 )
                wb[k * kWS + kSlice + lane] = v1[k];
        }
        // LDS operations of one wavefront complete in order; the fences keep
        // the compiler from moving the window reads above the stores.
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        if ( r + 1 < nrows )
            loadRow( r + 1 );
        float px = 0.0f, py = 0.0f, pz = 0.0f; // FP32 sum of this row's bond forces (mixed form)
        // the row's slots q = R.x + t through row base pointers, so the
        // unrolled slots address their words and table entries with
        // immediate offsets (no per-slot 64-bit address arithmetic)
        const uint64_t* __restrict__ mw_r = mwords + R.x;
        uint64_t* __restrict__ bits_r = bits + R.x;
        const float4* __restrict__ st_r = sten + R.x;
        const float4* __restrict__ sa_r = saux + R.x;
        const double* __restrict__ xv_r = sxiv + 4 * R.x;
        // XV: xi_y - u_i,y and xi_z - u_i,z, one value each for the row (the
        // slots of a lattice row differ in xi_x only)
        double ay = 0.0, az = 0.0;
        if constexpr ( XV )
        {
            ay = scalarLoad( xv_r + kXiY ) - ui1;
            az = scalarLoad( xv_r + kXiZ ) - ui2;
        }
        // The wave-uniform data of slot t: its intact word (fracture) or
        // existence word, its FP32 stencil entries and (XV) its FP64 offset
        // and |xi|^2 (one 32-byte record).  Loading slot t + 1's data while
        // slot t computes measured slower on the fracture plate (0.209 vs
        // 0.202 ms): the register copies between slots cost more scalar
        // issue than the hidden load latency saved.
        // (the 48-byte records of the other row passes, one scalar load for
        // all of a slot's values, measured 2.8 % slower here: the 32-byte FP64
        // record and the FP32 table entry are two loads but stay within one
        // scalar-cache line each)
        struct SlotK
        {
            uint64_t ww;
            float4 g, ga;
            double x0, x2d;
        };
        auto loadK = [&]( const int t ) __attribute__( ( always_inline ) ) {
            SlotK K{};
            K.ww = scalarLoad( mw_r + t );
            if constexpr ( PREC == kMixed )
            {
                K.g = scalarLoad4( st_r + t );
                K.ga = scalarLoad4( sa_r + t ); // (1/|xi|, |xi|^2, infl |xi|, infl |xi|^2)
                if constexpr ( XV )
                {
                    const double* xv = xv_r + 4 * t;
                    K.x0 = scalarLoad( xv + kXiX );
                    K.x2d = scalarLoad( xv + kXi2 );
                }
            }
            return K;
        };
        auto slot = [&]( const int t, const SlotK& K ) __attribute__( ( always_inline ) ) {
            const uint64_t ww = K.ww;
            // A slot with no live bond in the wave (the slots past a face of
            // the specimen, fully broken slots) adds nothing: the LPS passes
            // skip it (measured 8 % faster on the 128^3 LPS fracture cube; the
            // branch made the PMB plate pass 3 % slower, so PMB evaluates
            // and masks).  The FP64 form skips through its exec mask.
            if constexpr ( PREC == kMixed && kIsLps<MODEL> )
                if ( ww == 0 )
                    return;
            // The window always holds readable values; bonds that do not
            // exist are masked out below.
            const double* __restrict__ wr = wb + lane + t;
            const double uj0 = wr[0], uj1 = wr[kWS], uj2 = wr[2 * kWS];
            // candidates for the break test: intact, neither end no-fail
            // (LPS tests nofail(i) twice, force/CabanaPD_LPS.hpp:417-418)
            uint64_t cand = 0;
            if constexpr ( FRAC )
            {
                cand = ww & nfi_ok;
                if constexpr ( !kIsLps<MODEL> && GEN )
                    cand &= __ballot( wr[KNF * kWS] == 0.0 );
            }
            // Masks below are wave-uniform (ballots in converged code).
            uint64_t brkm = 0; // bonds that break in this pass
            if constexpr ( PREC == kF64 )
            {
                bool brk = false;
                if ( __builtin_amdgcn_inverse_ballot_w64( ww ) )
                {
                    const double vj = GEN ? wr[KV * kWS] : volu;
                    Bond64 b;
                    bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, wr[3 * kWS], wr[4 * kWS], wr[5 * kWS], uj0, uj1,
                                   uj2, b );
                    if constexpr ( FRAC )
                        brk = criticalStretch64( b.r, b.xi, bcc ) && __builtin_amdgcn_inverse_ballot_w64( cand );
                    if ( !brk )
                    {
                        double coeff;
                        if constexpr ( !kIsLps<MODEL> )
                            coeff = c * b.s * vj;
                        else
                        {
                            const double infl = infl1 ? 1.0 / b.xi : 1.0;
                            coeff = ( Ai + wr[( 3 + NX ) * kWS] + b.s * ( Bi + wr[( 4 + NX ) * kWS] ) ) * infl * b.xi * vj;
                        }
                        fx += coeff * b.rx / b.r;
                        fy += coeff * b.ry / b.r;
                        fz += coeff * b.rz / b.r;
                    }
                }
                if constexpr ( FRAC )
                    brkm = __ballot( brk );
            }
            else
            {
                const float4 g = K.g;
                const float4 ga = K.ga;
                Bond32 b;
                double r0 = 0.0, r1 = 0.0, r2 = 0.0; // XV: the FP64 bond vector
                if constexpr ( XV )
                {
                    r0 = ( uj0 - ui0 ) + K.x0;
                    r1 = uj1 + ay;
                    r2 = uj2 + az;
                    mixedBond( r0, r1, r2, K.x2d, g.w, b );
                }
                else
                    linearized32( g.x, g.y, g.z, g.w, ga.x, ui0, ui1, ui2, uj0, uj1, uj2, b );
                if constexpr ( FRAC )
                {
                    // FP32 decision, except within bandK of s0 where the
                    // reference's FP64 test decides.  One compare first:
                    // only a wave with a candidate at or above s0 - 2 bandK
                    // (rare: bonds near breaking) takes the decision path.
                    if ( __ballot( b.s >= s0lo ) & cand )
                    {
                    const uint64_t inb = __ballot( fabsf( b.s - s0f ) <= bandK ) & cand;
                    brkm = __ballot( b.s > s0f ) & cand & ~inb;
                    if ( inb )
                    {
                        bool crit = false;
                        if ( __builtin_amdgcn_inverse_ballot_w64( inb ) )
                        {
                            // neighbor j = p + R.z + t: uniform base + lane; the
                            // reference positions of i and j are loaded here
                            // (rare path; the compiler barrier keeps the loads of
                            // x_i from being hoisted into registers held for the
                            // whole pass), the displacements are the registers
                            // and the window
                            __asm__ volatile( "" ::: "memory" );
                            const int64_t jo = p0 + R.z + t;
                            const double* __restrict__ xj = x + jo;
                            const double* __restrict__ xp = x + p0;
                            Bond64 e;
                            getDistance64( xp[lane], ( xp + n )[lane], ( xp + 2 * n )[lane], ui0, ui1, ui2, xj[lane],
                                           ( xj + n )[lane], ( xj + 2 * n )[lane], uj0, uj1, uj2, e );
                            crit = criticalStretch64( e.r, e.xi, bcc );
                        }
                        brkm |= __ballot( crit ) & inb;
                    }
                    }
                }
                // dead lanes (absent, broken or breaking bonds) add exactly
                // zero: their window values are finite, fr is selected to 0
                const bool live = __builtin_amdgcn_inverse_ballot_w64( ww & ~brkm );
                float fr;
                if constexpr ( !kIsLps<MODEL> )
                    fr = ( GEN ? ( c32 * static_cast<float>( wr[KV * kWS] ) ) * b.s : b.s ) * b.rinv;
                else
                {
                    const float coeff = static_cast<float>( Ai + wr[( 3 + NX ) * kWS] ) +
                                        b.s * static_cast<float>( Bi + wr[( 4 + NX ) * kWS] );
                    fr = ( GEN ? ( coeff * ga.z ) * static_cast<float>( wr[KV * kWS] ) : coeff * ga.z ) * b.rinv;
                }
                fr = live ? fr : 0.0f;
                if constexpr ( XV )
                {
                    const double frd = static_cast<double>( fr );
                    sx = fma( frd, r0, sx );
                    sy = fma( frd, r1, sy );
                    sz = fma( frd, r2, sz );
                }
                else
                {
                    px = fmaf( fr, b.rx, px );
                    py = fmaf( fr, b.ry, py );
                    pz = fmaf( fr, b.rz, pz );
                }
            }
            if constexpr ( FRAC )
            {
                // wave-uniform test first (rare: a bond of the wave breaks),
                // then one lane writes the word
                if ( __builtin_expect( brkm != 0, 0 ) )
                {
                    if ( lane == 0 )
                        bits_r[t] = ww & ~brkm;
                }
            }
        };
        // Mixed forms: the slots of a row (at most kRowMaxSlots) unrolled, so
        // the window reads take immediate offsets and the PMB pass fits 70
        // VGPRs without spills (7 % faster on the fracture plate, 2 % on the
        // LPS cube; unrolling the FP64 form measured slower)
        if constexpr ( PREC == kMixed )
        {
#pragma unroll
            for ( int t = 0; t < kRowMaxSlots; t++ )
                if ( t < R.y )
                    slot( t, loadK( t ) );
        }
        else
        {
            for ( int t = 0; t < R.y; t++ )
                slot( t, loadK( t ) );
        }
        if constexpr ( PREC == kMixed && !XV )
        {
            fx = fma( rscale, static_cast<double>( px ), fx );
            fy = fma( rscale, static_cast<double>( py ), fy );
            fz = fma( rscale, static_cast<double>( pz ), fz );
        }
    }
    if constexpr ( XV )
    {
        fx = fma( rscale, sx, fx );
        fy = fma( rscale, sy, fy );
        fz = fma( rscale, sz, fz );
    }
    if ( !valid )
        return;
    epilogue( A, p, w.epi, w.cur, fx, fy, fz );
}

// ---------------------------------------------------------------------------
// The row-window force pass for specimens with per-particle features: several
// materials (pair tables), TemperatureDependent (per-particle temperature) and
// ElasticPerfectlyPlastic (per-bond plastic stretch).  FEAT is a set of
// kFeat* bits.  The window channels are u, x (FP64 form), lpsA/lpsB (LPS),
// the volume, a tag (2 x material + no-fail flag) and the temperature; the
// plastic stretch of slot q sits at splas[(row0 + q) * 64 + lane].  Per bond
// the arithmetic and the order of the sums are those of the gather pass
// (kForce with the slot layout), so both forms give identical results.
// ---------------------------------------------------------------------------
enum : int
{
    kFeatMT = 1, // pair tables (multi-material)
    kFeatTH = 2, // temperature-dependent stretch
    kFeatEP = 4  // elastic-perfectly-plastic bonds
};

// Waves per SIMD the register allocator must leave room for: 6 for the mixed
// PMB instances with several materials or a plastic stretch (measured 8 % and
// 10 % faster on the fracture plate than at 4-5 waves); the temperature
// instances spill at 6 and stay unconstrained.
template <int MODEL, bool FRAC, int PREC, int FEAT>
__global__ __launch_bounds__( kBlock ) __attribute__( ( amdgpu_waves_per_eu( ( PREC == kMixed && !kIsLps<MODEL> && ( FEAT == kFeatMT || FEAT == kFeatEP ) ) ? 6 : 1, 8 ) ) ) void kForceRowsF( const PDArgs* __restrict__ specs,
                                                         const int32_t* __restrict__ tile_spec,
                                                         const int64_t* __restrict__ tile_slice,
                                                         int32_t epi, int32_t parity, int64_t launch,
                                                         int32_t accumulate )
{
#pragma clang fp contract( off )
    constexpr bool MT = ( FEAT & kFeatMT ) != 0;
    constexpr bool TH = ( FEAT & kFeatTH ) != 0 && !kIsLps<MODEL>;
    constexpr bool EP = ( FEAT & kFeatEP ) != 0 && !kIsLps<MODEL>;
    constexpr int kW = kSlice + kRowMaxSlots - 1; // window width
    constexpr int NX = ( PREC == kF64 ) ? 3 : 0;  // staged x components
    constexpr int NAB = kIsLps<MODEL> ? 2 : 0;    // staged lpsA, lpsB
    constexpr int KV = 3 + NX + NAB;              // volume
    constexpr int KT = KV + 1;                    // tag: 2 x material + no-fail flag
    constexpr bool TAG = MT || ( FRAC && !kIsLps<MODEL> );
    constexpr int KH = KT + ( TAG ? 1 : 0 );      // temperature
    constexpr int NC = KH + ( TH ? 1 : 0 );
    // one buffer per wave: the next row's window waits in registers until
    // the current row's reads are done
    __shared__ double win[kSlicesPerBlock][NC][kW];
    // several materials: the pair tables (c, bond_break_coeff, FP32 s0 and
    // band) of the block's spec, indexed 4 a + b
    __shared__ double tab_c[MT ? 16 : 1], tab_bcc[MT ? 16 : 1];
    __shared__ float tab_s0[MT ? 16 : 1], tab_band[MT ? 16 : 1];
    WaveCtx w;
    const bool active = waveSetup( specs, tile_spec, tile_slice, epi, parity, launch, w );
    if constexpr ( MT )
    {
        // every wave of a block belongs to one spec
        if ( threadIdx.x < 16 )
        {
            tab_c[threadIdx.x] = w.A->mat_c[threadIdx.x];
            tab_bcc[threadIdx.x] = w.A->mat_bcc[threadIdx.x];
            tab_s0[threadIdx.x] = w.A->mat_s0f[threadIdx.x];
            tab_band[threadIdx.x] = w.A->mat_bandK[threadIdx.x];
        }
        __syncthreads();
    }
    if ( !active )
        return; // waves are otherwise independent: no further block-level synchronisation
    const PDArgs& A = *w.A;
    const int wv = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t n = scalarLoad( &A.n );
    const int64_t p0 = w.slice * kSlice;
    const int64_t p = p0 + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : p0;
    const double* __restrict__ x = scalarLoad( &A.x );
    const double* __restrict__ u = scalarLoad( &A.U[w.cur] );
    const int4* __restrict__ rows = scalarLoad( &A.rows );
    double xi0 = 0.0, xi1 = 0.0, xi2 = 0.0;
    if constexpr ( PREC == kF64 )
    {
        xi0 = x[pc];
        xi1 = x[n + pc];
        xi2 = x[2 * n + pc];
    }
    const double ui0 = u[pc], ui1 = u[n + pc], ui2 = u[2 * n + pc];
    const int64_t row0 = scalarLoad( scalarLoad( &A.slice_ptr ) + w.slice );
    uint64_t* __restrict__ bits = scalarLoad( &A.intact ) + row0;
    const uint64_t* __restrict__ mwords = FRAC ? bits : scalarLoad( &A.exist ) + row0;
    const double* __restrict__ srec = scalarLoad( &A.stencil_rec );
    const uint8_t* __restrict__ nfp = scalarLoad( &A.nofail );
    const uint8_t* __restrict__ mtp = MT ? scalarLoad( &A.mtype ) : nullptr;
    const double* __restrict__ tmpp = TH ? scalarLoad( &A.temp ) : nullptr;
    double* __restrict__ spl = EP ? scalarLoad( &A.splas ) + row0 * kSlice + lane : nullptr;
    const bool nf_i = FRAC ? ( nfp[pc] != 0 ) : false;
    const int ti4 = MT ? 4 * static_cast<int>( mtp[pc] ) : 0;
    const double Ti = TH ? tmpp[pc] : 0.0;
    const double c = A.c;
    const double bcc = A.bcc;
    const float c32 = static_cast<float>( A.c );
    const float s0f = A.s0f;
    const float bandK = A.bandK;
    const int infl1 = A.influence;
    // uniform volume: c vol (PMB, one material) or vol (LPS) scales the sum
    // (as in the gather pass)
    const bool fold = A.vol_uniform != 0 && !( FRAC && A.any_nofail != 0 ) && !( MT && !kIsLps<MODEL> );
    const double rscale = ( PREC == kMixed && fold ) ? A.row_scale : 1.0;
    double Ai = 0.0, Bi = 0.0;
    if constexpr ( kIsLps<MODEL> )
    {
        Ai = A.lpsA[pc]; // theta_coeff * theta_i / m_i
        Bi = A.lpsB[pc]; // s_coeff / m_i
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if ( accumulate && valid )
    {
        fx = A.f[p];
        fy = A.f[n + p];
        fz = A.f[2 * n + p];
    }
    double sx = 0.0, sy = 0.0, sz = 0.0; // mixed form: the bond forces fr r summed in FP64

    const double* src[KV + 1];
    src[0] = u;
    src[1] = u + n;
    src[2] = u + 2 * n;
    if constexpr ( NX )
    {
        src[3] = x;
        src[4] = x + n;
        src[5] = x + 2 * n;
    }
    if constexpr ( NAB )
    {
        src[3 + NX] = scalarLoad( &A.lpsA );
        src[4 + NX] = scalarLoad( &A.lpsB );
    }
    src[KV] = scalarLoad( &A.vol );
    // row tables exist only for n < 2^28 (byte sizes and offsets fit 32 bits)
    const int32_t p032 = static_cast<int32_t>( p0 );
    const uint32_t nbytes = static_cast<uint32_t>( n ) * 8u;
    __amdgpu_buffer_rsrc_t rs[KV + 1];
#pragma unroll
    for ( int k = 0; k <= KV; k++ )
        rs[k] = bufferOf( src[k], nbytes );
    const __amdgpu_buffer_rsrc_t rnf = bufferOf( nfp, static_cast<uint32_t>( n ) );
    const __amdgpu_buffer_rsrc_t rmt = bufferOf( MT ? mtp : nfp, static_cast<uint32_t>( n ) );
    const __amdgpu_buffer_rsrc_t rth = bufferOf( TH ? tmpp : src[0], nbytes );
    const uint32_t lane8 = static_cast<uint32_t>( lane ) * 8u;
    double v0[NC], v1[NC];
#pragma unroll
    for ( int k = 0; k < NC; k++ )
        v1[k] = 0.0;
    auto loadHalf = [&]( double* dst, int32_t g ) {
        // g >= 0: lane offset + uniform start; g < 0: per-lane offsets, the
        // negative ones wrap out of range and read 0
        const uint32_t vo = g >= 0 ? lane8 : static_cast<uint32_t>( g + lane ) * 8u;
        const uint32_t so = g >= 0 ? static_cast<uint32_t>( g ) * 8u : 0u;
#pragma unroll
        for ( int k = 0; k <= KV; k++ )
            dst[k] = bufferLoad64( rs[k], vo, so );
        if constexpr ( TAG )
        {
            int tag = 0;
            if constexpr ( FRAC && !kIsLps<MODEL> )
                tag = bufferLoad8( rnf, vo >> 3, so >> 3 ) != 0 ? 1 : 0;
            if constexpr ( MT )
                tag += 2 * static_cast<int>( bufferLoad8( rmt, vo >> 3, so >> 3 ) );
            dst[KT] = static_cast<double>( tag );
        }
        if constexpr ( TH )
            dst[KH] = bufferLoad64( rth, vo, so );
    };
    auto loadRow = [&]( int r ) {
        const int4 R = scalarLoad4( rows + r );
        const int32_t g = p032 + R.z; // uniform
        loadHalf( v0, g );
        if ( lane < R.y - 1 )
            loadHalf( v1, g + kSlice );
    };

    double* __restrict__ wb = &win[wv][0][0];
    const int nrows = scalarLoad( &A.nrows );
    loadRow( 0 );
    for ( int r = 0; r < nrows; r++ )
    {
        const int4 R = scalarLoad4( rows + r );
        // the previous row's window reads stay above this row's stores (a
        // lane's read of a position another lane overwrites)
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
#pragma unroll
        for ( int k = 0; k < NC; k++ )
        {
            wb[k * kW + lane] = v0[k];
            if ( lane < R.y - 1 )
                wb[k * kW + kSlice + lane] = v1[k];
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        if ( r + 1 < nrows )
            loadRow( r + 1 );
        auto slot = [&]( const int t ) __attribute__( ( always_inline ) ) {
            const int32_t q = R.x + t;
            const uint64_t ww = scalarLoad( mwords + q );
            const double* __restrict__ wr = wb + lane + t;
            const bool live = __builtin_amdgcn_inverse_ballot_w64( ww );
            int mj = 0;
            bool nf_j = false;
            if constexpr ( TAG )
            {
                const int tag = static_cast<int>( wr[KT * kW] );
                mj = tag >> 1;
                nf_j = ( tag & 1 ) != 0;
            }
            if constexpr ( kIsLps<MODEL> )
                nf_j = nf_i; // LPS tests nofail(i) twice (force/CabanaPD_LPS.hpp:417-418)
            // ElasticPerfectlyPlastic: the slot's plastic stretch is loaded
            // up front (every lane has an entry), so its latency overlaps the
            // bond arithmetic; it is written back only where it changes
            double spv = 0.0;
            if constexpr ( EP )
                spv = spl[q * kSlice];
            if constexpr ( PREC == kF64 )
            {
                bool brk = false;
                if ( live )
                {
                    const double uj0 = wr[0], uj1 = wr[kW], uj2 = wr[2 * kW];
                    const double vj = wr[KV * kW];
                    const double tj = TH ? wr[KH * kW] : 0.0;
                    Bond64 b;
                    bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, wr[3 * kW], wr[4 * kW], wr[5 * kW], uj0, uj1, uj2,
                                   b );
                    double temp_avg = 0.0;
                    if constexpr ( TH )
                    {
                        temp_avg = thermalAvg( A, Ti, tj );
                        b.s = b.s - ( A.alpha * temp_avg );
                    }
                    if constexpr ( FRAC )
                    {
                        const double bcc_ij = TH ? thermalBcc( A, temp_avg ) : ( MT ? tab_bcc[ti4 + mj] : bcc );
                        if ( criticalStretch64( b.r, b.xi, bcc_ij ) && !nf_i && !nf_j )
                            brk = true;
                    }
                    if ( !brk )
                    {
                        double coeff;
                        if constexpr ( !kIsLps<MODEL> )
                        {
                            const double sd = EP ? plasticStretchAt( spl + q * kSlice, spv, b.s, A.sY ) : b.s;
                            coeff = ( MT ? tab_c[ti4 + mj] : c ) * sd * vj;
                        }
                        else
                        {
                            const double infl = infl1 ? 1.0 / b.xi : 1.0;
                            coeff = ( Ai + wr[( 3 + NX ) * kW] + b.s * ( Bi + wr[( 4 + NX ) * kW] ) ) * infl * b.xi *
                                    vj;
                        }
                        fx += coeff * b.rx / b.r;
                        fy += coeff * b.ry / b.r;
                        fz += coeff * b.rz / b.r;
                    }
                }
                if constexpr ( FRAC )
                {
                    // wave-uniform test first (rare: a bond of the wave
                    // breaks), then one lane writes the word
                    const uint64_t m = __ballot( brk ) & ww;
                    if ( __builtin_expect( m != 0, 0 ) )
                    {
                        if ( lane == 0 )
                            bits[q] = ww & ~m;
                    }
                }
            }
            else
            {
                // every lane evaluates (window values are finite for every
                // lane); bonds that do not exist, are broken or break add
                // exactly zero.  r = xi + (u_j - u_i) from the slot's FP64
                // stencil record.
                const double* rr = srec + kRecStride * q;
                const double uj0 = wr[0], uj1 = wr[kW], uj2 = wr[2 * kW];
                const double r0 = ( uj0 - ui0 ) + scalarLoad( rr + kRecXiX );
                const double r1 = uj1 + ( scalarLoad( rr + kRecXiY ) - ui1 );
                const double r2 = uj2 + ( scalarLoad( rr + kRecXiZ ) - ui2 );
                const float xil = scalarLoad( reinterpret_cast<const float*>( rr + kRecF ) ); // |xi|
                Bond32 b;
                mixedBond( r0, r1, r2, scalarLoad( rr + kRecXi2 ), xil, b );
                double temp_avg = 0.0;
                float ath = 0.0f;
                const float s_mech = b.s;
                if constexpr ( TH )
                {
                    temp_avg = thermalAvg( A, Ti, wr[KH * kW] );
                    ath = static_cast<float>( A.alpha * temp_avg );
                    b.s = b.s - ath;
                }
                uint64_t brkm = 0; // bonds that break in this pass
                if constexpr ( FRAC )
                {
                    // candidates: intact, neither end no-fail; FP32 decision
                    // except within the pair's band of s0, where the
                    // reference's FP64 test decides
                    const uint64_t cand = ww & __ballot( !nf_i && !nf_j );
                    const float s0p = MT ? tab_s0[ti4 + mj] : s0f;
                    const float bandp = MT ? tab_band[ti4 + mj] : bandK;
                    const float band = TH ? fmaf( A.bandf, fabsf( s_mech ) + s0p + fabsf( ath ), bandp ) : bandp;
                    const uint64_t inb = __ballot( fabsf( b.s - s0p ) <= band ) & cand;
                    brkm = __ballot( b.s > s0p ) & cand & ~inb;
                    if ( inb )
                    {
                        bool crit = false;
                        if ( __builtin_amdgcn_inverse_ballot_w64( inb ) )
                        {
                            // rare path: the reference positions are loaded
                            // here (the compiler barrier keeps the loads of
                            // x_i from being hoisted into registers held for
                            // the whole pass)
                            __asm__ volatile( "" ::: "memory" );
                            const int64_t j = p + R.z + t;
                            Bond64 e;
                            getDistance64( x[p], x[n + p], x[2 * n + p], ui0, ui1, ui2, x[j], x[n + j], x[2 * n + j],
                                           uj0, uj1, uj2, e );
                            crit = criticalStretch64( e.r, e.xi,
                                                      TH ? thermalBcc( A, temp_avg ) : ( MT ? tab_bcc[ti4 + mj] : bcc ) );
                        }
                        brkm |= __ballot( crit ) & inb;
                    }
                }
                const bool ok = __builtin_amdgcn_inverse_ballot_w64( ww & ~brkm );
                float fr;
                if constexpr ( !kIsLps<MODEL> )
                {
                    // ElasticPerfectlyPlastic: the plastic stretch stays FP64
                    // and is updated on the bonds that carry force
                    float sd = b.s;
                    if constexpr ( EP )
                    {
                        if ( ok )
                            sd = static_cast<float>(
                                plasticStretchAt( spl + q * kSlice, spv, static_cast<double>( b.s ), A.sY ) );
                    }
                    const float cp = MT ? static_cast<float>( tab_c[ti4 + mj] ) : c32;
                    fr = ( fold ? sd : ( cp * static_cast<float>( wr[KV * kW] ) ) * sd ) * b.rinv;
                }
                else
                {
                    const float wl = infl1 ? 1.0f : xil; // infl |xi|
                    const float coeff = static_cast<float>( Ai + wr[( 3 + NX ) * kW] ) +
                                        b.s * static_cast<float>( Bi + wr[( 4 + NX ) * kW] );
                    fr = ( fold ? coeff * wl : ( coeff * wl ) * static_cast<float>( wr[KV * kW] ) ) * b.rinv;
                }
                fr = ok ? fr : 0.0f;
                const double frd = static_cast<double>( fr );
                sx = fma( frd, r0, sx );
                sy = fma( frd, r1, sy );
                sz = fma( frd, r2, sz );
                if constexpr ( FRAC )
                {
                    if ( __builtin_expect( brkm != 0, 0 ) )
                    {
                        if ( lane == 0 )
                            bits[q] = ww & ~brkm;
                    }
                }
            }
        };
        if constexpr ( PREC == kMixed )
        {
#pragma unroll
            for ( int t = 0; t < kRowMaxSlots; t++ )
                if ( t < R.y )
                    slot( t );
        }
        else
        {
            for ( int t = 0; t < R.y; t++ )
                slot( t );
        }
    }
    if constexpr ( PREC == kMixed )
    {
        fx = fma( rscale, sx, fx );
        fy = fma( rscale, sy, fy );
        fz = fma( rscale, sz, fz );
    }
    if ( !valid )
        return;
    epilogue( A, p, w.epi, w.cur, fx, fy, fz );
}

// ---------------------------------------------------------------------------
// LPS state pass: Force<LPS,*>::computeWeightedVolume + computeDilatation
// (force/CabanaPD_LPS.hpp:89-141, 287-378), fused: m_i and theta_i need only
// particle i's own bonds.  Also stores the per-particle force-coefficient
// terms lpsA = theta_coeff*theta/m and lpsB = s_coeff/m used by the force pass.
// ---------------------------------------------------------------------------
template <bool FRAC, int PREC>
__global__ __launch_bounds__( kBlock ) void kLpsState( const PDArgs* __restrict__ specs,
                                                       const int32_t* __restrict__ tile_spec,
                                                       const int64_t* __restrict__ tile_slice,
                                                       int32_t parity, int64_t launch,
                                                       int32_t do_m, int32_t do_theta )
{
#pragma clang fp contract( off )
    WaveCtx w;
    if ( !waveSetup( specs, tile_spec, tile_slice, 0, parity, launch, w ) )
        return;
    const PDArgs& A = *w.A;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t n = A.n;
    const int64_t p = w.slice * kSlice + lane;
    const bool valid = p < n;
    if ( !valid )
        return; // no cross-lane operations below
    const double* __restrict__ x = A.x;
    const double* __restrict__ u = A.U[w.cur];
    const double xi0 = x[p], xi1 = x[n + p], xi2 = x[2 * n + p];
    const double ui0 = u[p], ui1 = u[n + p], ui2 = u[2 * n + p];
    const int32_t mycnt = A.cnt[p];
    const int64_t base = A.slice_ptr[w.slice];
    const int32_t qend = slotEnd( A, static_cast<int32_t>( A.slice_ptr[w.slice + 1] - base ), mycnt );
    const int32_t* __restrict__ colp = A.col + base * kSlice + lane;
    const uint64_t* __restrict__ bits = A.intact + base;
    const double* __restrict__ vol = A.vol;
    const int infl1 = A.influence;
    // mixed form on a lattice: bond vectors from the stencil table (entry q
    // in the stencil-slot layout, the bond's stencil id in the compact one),
    // as in the row-window state pass
    const bool sten = ( PREC == kMixed ) && ( A.exist != nullptr || A.sid != nullptr );
    const uint8_t* __restrict__ sidp = A.sid ? A.sid + base * kSlice + lane : nullptr;

    double m_i;
    if ( do_m )
    {
        m_i = 0.0;
        for ( int32_t q = 0; q < qend; q++ )
        {
            if ( !slotBond( A, base, q, lane, mycnt ) )
                continue;
            if constexpr ( FRAC )
                if ( !( ( bits[q] >> lane ) & 1ull ) )
                    continue;
            const int64_t j = colp[q * kSlice];
            if constexpr ( PREC == kF64 )
            {
                Bond64 b;
                getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, x[j], x[n + j], x[2 * n + j], u[j], u[n + j],
                               u[2 * n + j], b );
                const double infl = infl1 ? 1.0 / b.xi : 1.0;
                m_i += infl * b.xi * b.xi * vol[j];
            }
            else
            {
                // infl |xi|^2: |xi| (influence 1) or |xi|^2
                float w2;
                if ( sten )
                {
                    // stencil entry: slot q (stencil-slot form) or the bond's id
                    const float xw = A.stencil[A.exist ? q : static_cast<int32_t>( sidp[q * kSlice] )].w;
                    w2 = infl1 ? xw : xw * xw;
                }
                else
                {
                    const double g0 = x[j] - xi0, g1 = x[n + j] - xi1, g2 = x[2 * n + j] - xi2;
                    const float x2f = static_cast<float>( fma( g0, g0, fma( g1, g1, g2 * g2 ) ) );
                    w2 = infl1 ? x2f * __builtin_amdgcn_rsqf( x2f ) : x2f;
                }
                m_i += static_cast<double>( w2 * static_cast<float>( vol[j] ) );
            }
        }
        A.m[p] = m_i;
    }
    else
        m_i = A.m[p];
    if ( !do_theta )
        return;
    double th = 0.0;
    const bool mpos = FRAC ? ( m_i > 0 ) : true;
    if ( mpos )
    {
        for ( int32_t q = 0; q < qend; q++ )
        {
            if ( !slotBond( A, base, q, lane, mycnt ) )
                continue;
            if constexpr ( FRAC )
                if ( !( ( bits[q] >> lane ) & 1ull ) )
                    continue;
            const int64_t j = colp[q * kSlice];
            if constexpr ( PREC == kF64 )
            {
                Bond64 b;
                getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, x[j], x[n + j], x[2 * n + j], u[j], u[n + j],
                               u[2 * n + j], b );
                const double infl = infl1 ? 1.0 / b.xi : 1.0;
                const double theta_b = infl * b.s * b.xi * b.xi * vol[j];
                th += 3.0 * theta_b / m_i;
            }
            else
            {
                Bond32 b;
                float w2;
                if ( sten )
                {
                    // r = xi + (u_j - u_i) with the slot's FP64 stencil offset,
                    // as in the row-window state pass
                    const int32_t id = A.exist ? q : static_cast<int32_t>( sidp[q * kSlice] );
                    const float xw = A.stencil[id].w;
                    const double* xv = A.stencil_xiv + 4 * id;
                    mixedBond( ( u[j] - ui0 ) + xv[kXiX], u[n + j] + ( xv[kXiY] - ui1 ), u[2 * n + j] + ( xv[kXiZ] - ui2 ),
                               xv[kXi2], xw, b );
                    w2 = infl1 ? xw : xw * xw;
                }
                else
                {
                    const double g0 = x[j] - xi0, g1 = x[n + j] - xi1, g2 = x[2 * n + j] - xi2;
                    const double r0 = ( u[j] - ui0 ) + g0, r1 = ( u[n + j] - ui1 ) + g1, r2 = ( u[2 * n + j] - ui2 ) + g2;
                    const double xi2d = fma( g0, g0, fma( g1, g1, g2 * g2 ) );
                    const float x2f = static_cast<float>( xi2d );
                    const float xil = x2f * __builtin_amdgcn_rsqf( x2f );
                    mixedBond( r0, r1, r2, xi2d, xil, b );
                    w2 = infl1 ? xil : x2f;
                }
                th += static_cast<double>( w2 * b.s * static_cast<float>( vol[j] ) );
            }
        }
        if constexpr ( PREC == kMixed )
            if ( mycnt > 0 )
                th = 3.0 * th / m_i;
    }
    A.theta[p] = th;
    // ForceCoeffTag per-particle subexpressions (force_models/CabanaPD_LPS.hpp:131-132),
    // with the particle's own material coefficients in a multi-material specimen
    const double tc = A.mtype ? A.mat_tc[A.mtype[p]] : A.theta_coeff;
    const double sc = A.mtype ? A.mat_sc[A.mtype[p]] : A.s_coeff;
    A.lpsA[p] = tc * th / m_i;
    A.lpsB[p] = sc / m_i;
}

// ---------------------------------------------------------------------------
// The LPS state pass with neighbor data staged in LDS row windows (stencil-slot
// form on lattice rows), as kForceRows does for the force pass.  Same bonds,
// order and expressions as kLpsState, so the results are identical; the FP64
// form walks the rows twice (m_i must be complete before each bond's share of
// theta_i is divided by it), the mixed form once.
// ---------------------------------------------------------------------------
template <bool FRAC, int PREC, bool GEN>
__global__ __launch_bounds__( kBlock ) void kLpsStateRows( const PDArgs* __restrict__ specs,
                                                           const int32_t* __restrict__ tile_spec,
                                                           const int64_t* __restrict__ tile_slice,
                                                           int32_t parity, int64_t launch,
                                                           int32_t do_m, int32_t do_theta )
{
#pragma clang fp contract( off )
    constexpr int kW = kSlice + kRowMaxSlots - 1;
    constexpr int NX = ( PREC == kF64 ) ? 3 : 0; // staged x components
    constexpr int KV = 3 + NX;                   // staged vol (GEN)
    // mixed form: the bond vector r = xi + (u_j - u_i) in FP64 from the slot's
    // FP64 stencil offset xi; the windows hold u only
    constexpr int NC = KV + ( GEN ? 1 : 0 );
    __shared__ double win[kSlicesPerBlock][2][NC][kW];
    WaveCtx w;
    if ( !waveSetup( specs, tile_spec, tile_slice, 0, parity, launch, w ) )
        return; // waves are independent: no block-level synchronisation
    const PDArgs& A = *w.A;
    const int wv = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t n = scalarLoad( &A.n );
    const int64_t p0 = w.slice * kSlice;
    const int64_t p = p0 + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : p0;
    const double* __restrict__ x = scalarLoad( &A.x );
    const double* __restrict__ u = scalarLoad( &A.U[w.cur] );
    const int4* __restrict__ rows = scalarLoad( &A.rows );
    const double xi0 = x[pc], xi1 = x[n + pc], xi2 = x[2 * n + pc];
    const double ui0 = u[pc], ui1 = u[n + pc], ui2 = u[2 * n + pc];
    // per-slot lane masks (intact words, a subset of the existence words,
    // or the existence words without fracture) and the stencil tables, all
    // wave-uniform: scalar loads
    const int64_t row0 = scalarLoad( scalarLoad( &A.slice_ptr ) + w.slice );
    const uint64_t* __restrict__ mwords = scalarLoad( FRAC ? &A.intact : &A.exist ) + row0;
    const double* __restrict__ srec = scalarLoad( &A.stencil_rec );
    const int infl1 = A.influence;
    const double volu = A.vol_value;
    const float vol32 = static_cast<float>( A.vol_value );

    const double* src[KV];
    src[0] = u;
    src[1] = u + n;
    src[2] = u + 2 * n;
    if constexpr ( NX )
    {
        src[3] = x;
        src[4] = x + n;
        src[5] = x + 2 * n;
    }
    const double* __restrict__ volp = scalarLoad( &A.vol );
    const int32_t p032 = static_cast<int32_t>( p0 );
    double v0[NC], v1[NC];
#pragma unroll
    for ( int k = 0; k < NC; k++ )
        v1[k] = 0.0;
    // Row windows through raw buffer loads, as in kForceRows (n < 2^28).
    const uint32_t nbytes = static_cast<uint32_t>( n ) * 8u;
    __amdgpu_buffer_rsrc_t rs[NC];
#pragma unroll
    for ( int k = 0; k < KV; k++ )
        rs[k] = bufferOf( src[k], nbytes );
    if constexpr ( GEN )
        rs[KV] = bufferOf( volp, nbytes );
    const uint32_t lane8 = static_cast<uint32_t>( lane ) * 8u;
    auto loadHalf = [&]( double* dst, int32_t g ) {
        const uint32_t vo = g >= 0 ? lane8 : static_cast<uint32_t>( g + lane ) * 8u;
        const uint32_t so = g >= 0 ? static_cast<uint32_t>( g ) * 8u : 0u;
#pragma unroll
        for ( int k = 0; k < NC; k++ )
            dst[k] = bufferLoad64( rs[k], vo, so );
    };
    auto loadRow = [&]( int r ) {
        const int4 R = scalarLoad4( rows + r );
        const int32_t g = p032 + R.z; // uniform
        loadHalf( v0, g );
        if ( lane < R.y - 1 )
            loadHalf( v1, g + kSlice );
    };
    // Walk every bond slot in slot order; bond(live, wr, q) gets the window
    // position of the neighbor's staged values.
    auto walk = [&]( auto&& bond ) {
        const int nrows = scalarLoad( &A.nrows );
        loadRow( 0 );
        for ( int r = 0; r < nrows; r++ )
        {
            const int4 R = scalarLoad4( rows + r );
            double* __restrict__ wb = &win[wv][r & 1][0][0];
#pragma unroll
            for ( int k = 0; k < NC; k++ )
            {
                wb[k * kW + lane] = v0[k];
                if ( lane < R.y - 1 )
                    wb[k * kW + kSlice + lane] = v1[k];
            }
            __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
            if ( r + 1 < nrows )
                loadRow( r + 1 );
            // mixed form: the row's slots unrolled (immediate window offsets);
            // xi_y - u_i,y and xi_z - u_i,z once per row (the slots of a
            // lattice row differ in xi_x only)
            if constexpr ( PREC == kMixed )
            {
                const double* rec_r = srec + kRecStride * R.x;
                const double ay = scalarLoad( rec_r + kRecXiY ) - ui1;
                const double az = scalarLoad( rec_r + kRecXiZ ) - ui2;
#pragma unroll
                for ( int t = 0; t < kRowMaxSlots; t++ )
                    if ( t < R.y )
                    {
                        const int32_t q = R.x + t;
                        const bool live = __builtin_amdgcn_inverse_ballot_w64( scalarLoad( mwords + q ) );
                        bond( live, wb + lane + t, q, ay, az );
                    }
            }
            else
            {
                for ( int t = 0; t < R.y; t++ )
                {
                    const int32_t q = R.x + t;
                    const bool live = __builtin_amdgcn_inverse_ballot_w64( scalarLoad( mwords + q ) );
                    bond( live, wb + lane + t, q, 0.0, 0.0 );
                }
            }
        }
    };

    double m_i = 0.0;
    double th = 0.0;
    if constexpr ( PREC == kF64 )
    {
        if ( do_m )
        {
            walk( [&]( bool live, const double* wr, int32_t, double, double ) {
                if ( live )
                {
                    Bond64 b;
                    getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, wr[3 * kW], wr[4 * kW], wr[5 * kW], wr[0], wr[kW],
                                   wr[2 * kW], b );
                    const double infl = infl1 ? 1.0 / b.xi : 1.0;
                    m_i += infl * b.xi * b.xi * ( GEN ? wr[KV * kW] : volu );
                }
            } );
        }
        else if ( valid )
            m_i = A.m[p];
        const bool mpos = FRAC ? ( m_i > 0 ) : true;
        if ( do_theta )
        {
            walk( [&]( bool live, const double* wr, int32_t, double, double ) {
                if ( live && mpos )
                {
                    Bond64 b;
                    getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, wr[3 * kW], wr[4 * kW], wr[5 * kW], wr[0], wr[kW],
                                   wr[2 * kW], b );
                    const double infl = infl1 ? 1.0 / b.xi : 1.0;
                    const double theta_b = infl * b.s * b.xi * b.xi * ( GEN ? wr[KV * kW] : volu );
                    th += 3.0 * theta_b / m_i;
                }
            } );
        }
    }
    else
    {
        if ( !do_m && valid )
            m_i = A.m[p];
        // m_i for the theta condition is only known after the walk: sum both,
        // apply the condition at the end (the sums are the same per-bond terms)
        walk( [&]( bool live, const double* wr, int32_t q, double ay, double az ) {
            if ( live )
            {
                // the slot's record: xi_x, |xi|^2, |xi|, infl |xi|^2
                const double* rr = srec + kRecStride * q;
                const float* rf = reinterpret_cast<const float*>( rr + kRecF );
                const float w2 = scalarLoad( rf + 2 ); // infl |xi|^2
                const float vj32 = GEN ? static_cast<float>( wr[KV * kW] ) : vol32;
                if ( do_m )
                    m_i += static_cast<double>( w2 * vj32 );
                if ( do_theta )
                {
                    Bond32 b;
                    mixedBond( ( wr[0] - ui0 ) + scalarLoad( rr + kRecXiX ), wr[kW] + ay, wr[2 * kW] + az,
                               scalarLoad( rr + kRecXi2 ), scalarLoad( rf ), b );
                    th += static_cast<double>( w2 * b.s * vj32 );
                }
            }
        } );
        if ( do_theta )
        {
            const bool mpos = FRAC ? ( m_i > 0 ) : true;
            if ( !mpos )
                th = 0.0;
            else if ( valid && A.cnt[p] > 0 )
                th = 3.0 * th / m_i;
        }
    }
    if ( !valid )
        return;
    if ( do_m )
        A.m[p] = m_i;
    if ( !do_theta )
        return;
    A.theta[p] = th;
    // ForceCoeffTag per-particle subexpressions (force_models/CabanaPD_LPS.hpp:131-132),
    // with the particle's own material coefficients in a multi-material specimen
    const double tc = A.mtype ? A.mat_tc[A.mtype[p]] : A.theta_coeff;
    const double sc = A.mtype ? A.mat_sc[A.mtype[p]] : A.s_coeff;
    A.lpsA[p] = tc * th / m_i;
    A.lpsB[p] = sc / m_i;
}

// ---------------------------------------------------------------------------
// Energy / damage: Force<*>::computeEnergyFull (force/CabanaPD_PMB.hpp:127-155,
// 289-339; force/CabanaPD_LPS.hpp:183-219, 449-504).  FP64 for every form.
// Per-particle W and phi; per-block partial sums (Phi, D) reduced on the host
// in block order, so totals are reproducible run to run.
// ---------------------------------------------------------------------------
template <int MODEL, bool FRAC>
__global__ __launch_bounds__( kBlock ) void kEnergy( const PDArgs* __restrict__ specs, int32_t parity,
                                                     const double* __restrict__ mvals,
                                                     const double* __restrict__ thetavals,
                                                     double* __restrict__ partial )
{
#pragma clang fp contract( off )
    __shared__ double red[2][kBlock / kSlice];
    const PDArgs& A = *specs;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int wave = threadIdx.x / kSlice;
    const int64_t slice = blockIdx.x * (int64_t)kSlicesPerBlock + wave;
    const int64_t n = A.n;
    const int64_t p = slice * kSlice + lane;
    double Phi = 0.0, D = 0.0;
    if ( slice < A.nslices && p < n )
    {
        const double* __restrict__ x = A.x;
        const double* __restrict__ u = A.U[parity];
        const double xi0 = x[p], xi1 = x[n + p], xi2 = x[2 * n + p];
        const double ui0 = u[p], ui1 = u[n + p], ui2 = u[2 * n + p];
        const int32_t mycnt = A.cnt[p];
        const int64_t base = A.slice_ptr[slice];
        const int32_t qend = slotEnd( A, static_cast<int32_t>( A.slice_ptr[slice + 1] - base ), mycnt );
        const int32_t* __restrict__ colp = A.col + base * kSlice + lane;
        const uint64_t* __restrict__ bits = A.intact + base;
        const double* __restrict__ vol = A.vol;
        const double c = A.c;
        const uint8_t* __restrict__ mt = A.mtype; // multi-material: pair / own-material constants
        const int ti4 = mt ? 4 * static_cast<int>( mt[p] ) : 0;
        const double tc_i = mt ? A.mat_tc[mt[p]] : A.theta_coeff;
        const double sc_i = mt ? A.mat_sc[mt[p]] : A.s_coeff;
        const double* __restrict__ tmp = kIsLps<MODEL> ? nullptr : A.temp;
        const double* __restrict__ spl = kIsLps<MODEL> ? nullptr : A.splas;
        const double Ti = tmp ? tmp[p] : 0.0;
        double num_bonds = 0.0;
        double m_i = 0.0, theta_i = 0.0;
        if constexpr ( kIsLps<MODEL> )
        {
            m_i = mvals[p];
            theta_i = thetavals[p];
            if constexpr ( FRAC )
            {
                for ( int32_t q = 0; q < qend; q++ )
                    num_bonds += static_cast<double>( ( bits[q] >> lane ) & 1ull ); // 0 on empty slots
            }
            else
                num_bonds = static_cast<double>( mycnt );
        }
        double W = 0.0, phi_i = 0.0, vol_H_i = 0.0;
        const double vol_i = vol[p];
        for ( int32_t q = 0; q < qend; q++ )
        {
            if ( !slotBond( A, base, q, lane, mycnt ) )
                continue;
            const int64_t j = colp[q * kSlice];
            Bond64 b;
            bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, x[j], x[n + j], x[2 * n + j], u[j], u[n + j],
                           u[2 * n + j], b );
            double wb;
            if constexpr ( kIsLps<MODEL> )
            {
                const double infl = A.influence ? 1.0 / b.xi : 1.0;
                wb = 1.0 / num_bonds * 0.5 * tc_i / 3.0 * ( theta_i * theta_i ) +
                     0.5 * ( sc_i / m_i ) * infl * b.s * b.s * b.xi * b.xi * vol[j];
            }
            else
            {
                if ( tmp )
                    b.s = b.s - ( A.alpha * thermalAvg( A, Ti, tmp[j] ) );
                const double cc = mt ? A.mat_c[ti4 + mt[j]] : c;
                if ( spl )
                    wb = 0.25 * cc * plasticEnergyTerm( b.s, spl[( base + q ) * kSlice + lane], A.sY ) * b.xi * vol[j];
                else
                    wb = 0.25 * cc * b.s * b.s * b.xi * vol[j];
            }
            if constexpr ( FRAC )
            {
                const double mu = static_cast<double>( ( bits[q] >> lane ) & 1ull );
                wb = mu * wb;
                W += wb;
                phi_i += mu * vol[j];
                vol_H_i += vol[j];
            }
            else
            {
                W += wb;
                Phi += wb * vol_i;
            }
        }
        A.W[p] = W;
        if constexpr ( FRAC )
        {
            Phi = W * vol_i;
            const double ph = 1 - phi_i / vol_H_i;
            A.phi[p] = ph;
            D = ph;
        }
    }
    for ( int off = 32; off > 0; off >>= 1 )
    {
        Phi += __shfl_xor( Phi, off );
        D += __shfl_xor( D, off );
    }
    if ( lane == 0 )
    {
        red[0][wave] = Phi;
        red[1][wave] = D;
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        double a = 0.0, d = 0.0;
        for ( int k = 0; k < kSlicesPerBlock; k++ )
        {
            a += red[0][k];
            d += red[1][k];
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = d;
    }
}

// ---------------------------------------------------------------------------
// Stress: Force<*>::computeStressFull (force/CabanaPD_PMB.hpp:157-202,
// 341-401, 492-536; force/CabanaPD_LPS.hpp:221-267, 506-572, 677-723) after
// computeStress's reset (CabanaPD_Force.hpp:221-236).  Per bond, f is half the
// bond force (PMB: 0.5 * (c s V_j), LPS: the force coefficient times 0.5; the
// fracture forms multiply by mu, LPS fracture visits unbroken bonds only) and
// stress(i, a, b) += f_a xi_b with xi the reference bond.  FP64 in every
// precision; one lane per particle, nine accumulators, written [9][n].
// ---------------------------------------------------------------------------
template <int MODEL, bool FRAC>
__global__ __launch_bounds__( kBlock ) void kStress( const PDArgs* __restrict__ specs, int32_t parity )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int wave = threadIdx.x / kSlice;
    const int64_t slice = blockIdx.x * (int64_t)kSlicesPerBlock + wave;
    const int64_t n = A.n;
    const int64_t p = slice * kSlice + lane;
    if ( slice >= A.nslices || p >= n )
        return;
    const double* __restrict__ x = A.x;
    const double* __restrict__ u = A.U[parity];
    const double xi0 = x[p], xi1 = x[n + p], xi2 = x[2 * n + p];
    const double ui0 = u[p], ui1 = u[n + p], ui2 = u[2 * n + p];
    const int32_t mycnt = A.cnt[p];
    const int64_t base = A.slice_ptr[slice];
    const int32_t qend = slotEnd( A, static_cast<int32_t>( A.slice_ptr[slice + 1] - base ), mycnt );
    const int32_t* __restrict__ colp = A.col + base * kSlice + lane;
    const uint64_t* __restrict__ bits = A.intact + base;
    const double* __restrict__ vol = A.vol;
    const double c = A.c;
    const uint8_t* __restrict__ mt = A.mtype; // multi-material: pair constants
    const int ti4 = mt ? 4 * static_cast<int>( mt[p] ) : 0;
    const double* __restrict__ tmp = kIsLps<MODEL> ? nullptr : A.temp;
    double* __restrict__ spl = kIsLps<MODEL> ? nullptr : A.splas;
    const double Ti = tmp ? tmp[p] : 0.0;
    double Ai = 0.0, Bi = 0.0;
    if constexpr ( kIsLps<MODEL> )
    {
        Ai = A.lpsA[p];
        Bi = A.lpsB[p];
    }
    double s00 = 0.0, s01 = 0.0, s02 = 0.0, s10 = 0.0, s11 = 0.0, s12 = 0.0, s20 = 0.0, s21 = 0.0, s22 = 0.0;
    for ( int32_t q = 0; q < qend; q++ )
    {
        if ( !slotBond( A, base, q, lane, mycnt ) )
            continue;
        double mu = 1.0;
        if constexpr ( FRAC )
        {
            mu = static_cast<double>( ( bits[q] >> lane ) & 1ull );
            if ( kIsLps<MODEL> && !( mu > 0 ) )
                continue;
        }
        const int64_t j = colp[q * kSlice];
        const double xj0 = x[j], xj1 = x[n + j], xj2 = x[2 * n + j];
        Bond64 b;
        bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, xj0, xj1, xj2, u[j], u[n + j], u[2 * n + j], b );
        const double xi_x = xj0 - xi0;
        const double xi_y = xj1 - xi1;
        const double xi_z = xj2 - xi2;
        double coeff;
        if constexpr ( kIsLps<MODEL> )
        {
            const double infl = A.influence ? 1.0 / b.xi : 1.0;
            coeff = ( Ai + A.lpsA[j] + b.s * ( Bi + A.lpsB[j] ) ) * infl * b.xi * vol[j];
            coeff *= 0.5;
        }
        else
        {
            if ( tmp )
                b.s = b.s - ( A.alpha * thermalAvg( A, Ti, tmp[j] ) );
            // EPP: the reference's stress pass updates s_p of every bond (:373-376)
            const double sd = spl ? plasticStretch( spl + ( base + q ) * kSlice + lane, b.s, A.sY ) : b.s;
            coeff = 0.5 * ( ( mt ? A.mat_c[ti4 + mt[j]] : c ) * sd * vol[j] );
        }
        double fx, fy, fz;
        if constexpr ( FRAC )
        {
            fx = mu * coeff * b.rx / b.r;
            fy = mu * coeff * b.ry / b.r;
            fz = mu * coeff * b.rz / b.r;
        }
        else
        {
            fx = coeff * b.rx / b.r;
            fy = coeff * b.ry / b.r;
            fz = coeff * b.rz / b.r;
        }
        s00 += fx * xi_x;
        s11 += fy * xi_y;
        s22 += fz * xi_z;
        s01 += fx * xi_y;
        s10 += fy * xi_x;
        s02 += fx * xi_z;
        s20 += fz * xi_x;
        s12 += fy * xi_z;
        s21 += fz * xi_y;
    }
    double* __restrict__ st = A.stress;
    st[p] = s00;
    st[n + p] = s01;
    st[2 * n + p] = s02;
    st[3 * n + p] = s10;
    st[4 * n + p] = s11;
    st[5 * n + p] = s12;
    st[6 * n + p] = s20;
    st[7 * n + p] = s21;
    st[8 * n + p] = s22;
}

// ---------------------------------------------------------------------------
// Output passes on lattice-row windows: the energy and stress passes above
// for the stencil-slot layout with a row table, with no per-particle material
// type, temperature or plastic stretch (those keep the gathers).  Each
// wavefront stages the neighbour windows of its slice's lattice rows in LDS
// (u, x, then NE further per-particle arrays, then vol when volumes differ)
// with raw buffer loads, as the row-window force pass does, and visits the
// slots in slot order, so every per-particle sum takes the gathers' terms in
// the gathers' order: results are bit-identical to kEnergy / kStress.
// ---------------------------------------------------------------------------
template <int NC>
struct RowWalk64
{
    static constexpr int kW = kSlice + kRowMaxSlots - 1; // window width
    __amdgpu_buffer_rsrc_t rs[NC];
    const int4* __restrict__ rows;
    int nrows;
    int32_t p032; // first particle of the slice
    int lane;

    // bond(q, wr, g) for every slot q of the slice in slot order; wr is the
    // lane's window position (channel k at wr[k * kW]) and g + lane its
    // neighbour's id.  win: this wave's [2][NC][kW] LDS buffers.
    template <class Bond>
    __device__ __forceinline__ void walk( double* __restrict__ win, Bond&& bond ) const
    {
        double v0[NC], v1[NC];
#pragma unroll
        for ( int k = 0; k < NC; k++ )
            v1[k] = 0.0;
        const uint32_t lane8 = static_cast<uint32_t>( lane ) * 8u;
        auto loadHalf = [&]( double* dst, int32_t g ) {
            // g < 0: per-lane offsets, the negative ones wrap out of range and read 0
            const uint32_t vo = g >= 0 ? lane8 : static_cast<uint32_t>( g + lane ) * 8u;
            const uint32_t so = g >= 0 ? static_cast<uint32_t>( g ) * 8u : 0u;
#pragma unroll
            for ( int k = 0; k < NC; k++ )
                dst[k] = bufferLoad64( rs[k], vo, so );
        };
        auto loadRow = [&]( int r ) {
            const int4 R = scalarLoad4( rows + r );
            const int32_t g = p032 + R.z; // uniform
            loadHalf( v0, g );
            if ( lane < R.y - 1 )
                loadHalf( v1, g + kSlice );
        };
        loadRow( 0 );
        for ( int r = 0; r < nrows; r++ )
        {
            const int4 R = scalarLoad4( rows + r );
            double* __restrict__ wb = win + ( r & 1 ) * NC * kW;
#pragma unroll
            for ( int k = 0; k < NC; k++ )
            {
                wb[k * kW + lane] = v0[k];
                if ( lane < R.y - 1 )
                    wb[k * kW + kSlice + lane] = v1[k];
            }
            // LDS operations of one wavefront complete in order; the fences
            // keep the compiler from moving the window reads above the stores
            __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
            if ( r + 1 < nrows )
                loadRow( r + 1 );
            for ( int t = 0; t < R.y; t++ )
                bond( R.x + t, wb + lane + t, p032 + R.z + t );
        }
    }
};

// Set up the walker of slice `slice` over channels src[0 .. NC) (n values each).
template <int NC>
__device__ __forceinline__ RowWalk64<NC> rowWalk( const PDArgs& A, int64_t slice, int lane, const double* const* src )
{
    RowWalk64<NC> w;
    const int64_t n = scalarLoad( &A.n );
    // row tables exist only for n < 2^28 (byte sizes and offsets fit 32 bits)
    const uint32_t nbytes = static_cast<uint32_t>( n ) * 8u;
#pragma unroll
    for ( int k = 0; k < NC; k++ )
        w.rs[k] = bufferOf( src[k], nbytes );
    w.rows = scalarLoad( &A.rows );
    w.nrows = scalarLoad( &A.nrows );
    w.p032 = static_cast<int32_t>( slice * kSlice );
    w.lane = lane;
    return w;
}

// FEAT (kFeat* bits): several materials (the neighbour's type read at its
// id), a temperature (a staged channel), a plastic stretch (per slot).
template <int MODEL, bool FRAC, bool GEN, int FEAT = 0>
__global__ __launch_bounds__( kBlock ) void kEnergyRows( const PDArgs* __restrict__ specs, int32_t parity,
                                                         const double* __restrict__ mvals,
                                                         const double* __restrict__ thetavals,
                                                         double* __restrict__ partial )
{
#pragma clang fp contract( off )
    constexpr bool MT = ( FEAT & kFeatMT ) != 0;
    constexpr bool TH = ( FEAT & kFeatTH ) != 0 && !kIsLps<MODEL>;
    constexpr bool EP = ( FEAT & kFeatEP ) != 0 && !kIsLps<MODEL>;
    constexpr int KT = 6 + ( GEN ? 1 : 0 ); // temperature channel
    constexpr int NC = KT + ( TH ? 1 : 0 );
    constexpr int kW = RowWalk64<NC>::kW;
    __shared__ double win[kSlicesPerBlock][2][NC][kW];
    __shared__ double red[2][kSlicesPerBlock];
    const PDArgs& A = *specs;
    const int wv = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int lane = threadIdx.x & ( kSlice - 1 );
    // the partial sums of logical block t, as in kEnergy
    const int64_t t = xcdRemap( blockIdx.x, gridDim.x );
    const int64_t slice = t * kSlicesPerBlock + wv;
    double Phi = 0.0, D = 0.0;
    if ( slice < A.nslices )
    {
        const int64_t n = scalarLoad( &A.n );
        const int64_t p0 = slice * kSlice;
        const int64_t p = p0 + lane;
        const bool valid = p < n;
        const int64_t pc = valid ? p : p0;
        const double* __restrict__ x = scalarLoad( &A.x );
        const double* __restrict__ u = scalarLoad( &A.U[parity] );
        const double* __restrict__ volp = scalarLoad( &A.vol );
        const double* src[NC];
        src[0] = u;
        src[1] = u + n;
        src[2] = u + 2 * n;
        src[3] = x;
        src[4] = x + n;
        src[5] = x + 2 * n;
        if constexpr ( GEN )
            src[6] = volp;
        if constexpr ( TH )
            src[KT] = scalarLoad( &A.temp );
        const RowWalk64<NC> rw = rowWalk<NC>( A, slice, lane, src );
        const double xi0 = x[pc], xi1 = x[n + pc], xi2 = x[2 * n + pc];
        const double ui0 = u[pc], ui1 = u[n + pc], ui2 = u[2 * n + pc];
        // per-slot lane masks: existence (every bond, broken or not) and intact
        const int64_t row0 = scalarLoad( scalarLoad( &A.slice_ptr ) + slice );
        const int32_t nslot = static_cast<int32_t>( scalarLoad( scalarLoad( &A.slice_ptr ) + slice + 1 ) - row0 );
        const uint64_t* __restrict__ ew = scalarLoad( &A.exist ) + row0;
        const uint64_t* __restrict__ iw = scalarLoad( &A.intact ) + row0;
        const uint8_t* __restrict__ mtp = MT ? scalarLoad( &A.mtype ) : nullptr;
        const double* __restrict__ spl = EP ? scalarLoad( &A.splas ) + row0 * kSlice + lane : nullptr;
        const int ti4 = MT ? 4 * static_cast<int>( mtp[pc] ) : 0;
        const double Ti = TH ? A.temp[pc] : 0.0;
        const double c = A.c;
        const double tc_i = MT ? A.mat_tc[mtp[pc]] : A.theta_coeff;
        const double sc_i = MT ? A.mat_sc[mtp[pc]] : A.s_coeff;
        const double volu = A.vol_value;
        const int infl1 = A.influence;
        double num_bonds = 0.0;
        double m_i = 0.0, theta_i = 0.0;
        if constexpr ( kIsLps<MODEL> )
        {
            m_i = mvals[pc];
            theta_i = thetavals[pc];
            if constexpr ( FRAC )
            {
                for ( int32_t q = 0; q < nslot; q++ )
                    num_bonds += static_cast<double>( ( scalarLoad( iw + q ) >> lane ) & 1ull ); // 0 on empty slots
            }
            else
                num_bonds = static_cast<double>( A.cnt[pc] );
        }
        const double vol_i = volp[pc];
        double W = 0.0, phi_i = 0.0, vol_H_i = 0.0;
        rw.walk( &win[wv][0][0][0], [&]( int32_t q, const double* __restrict__ wr, int32_t g ) {
            if ( !__builtin_amdgcn_inverse_ballot_w64( scalarLoad( ew + q ) ) )
                return;
            Bond64 b;
            bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, wr[3 * kW], wr[4 * kW], wr[5 * kW], wr[0], wr[kW],
                           wr[2 * kW], b );
            const double vj = GEN ? wr[6 * kW] : volu;
            double wb;
            if constexpr ( kIsLps<MODEL> )
            {
                const double infl = infl1 ? 1.0 / b.xi : 1.0;
                wb = 1.0 / num_bonds * 0.5 * tc_i / 3.0 * ( theta_i * theta_i ) +
                     0.5 * ( sc_i / m_i ) * infl * b.s * b.s * b.xi * b.xi * vj;
            }
            else
            {
                if constexpr ( TH )
                    b.s = b.s - ( A.alpha * thermalAvg( A, Ti, wr[KT * kW] ) );
                const double cc = MT ? A.mat_c[ti4 + mtp[g + lane]] : c;
                if constexpr ( EP )
                    wb = 0.25 * cc * plasticEnergyTerm( b.s, spl[q * kSlice], A.sY ) * b.xi * vj;
                else
                    wb = 0.25 * cc * b.s * b.s * b.xi * vj;
            }
            if constexpr ( FRAC )
            {
                const double mu = static_cast<double>( ( scalarLoad( iw + q ) >> lane ) & 1ull );
                wb = mu * wb;
                W += wb;
                phi_i += mu * vj;
                vol_H_i += vj;
            }
            else
            {
                W += wb;
                Phi += wb * vol_i;
            }
        } );
        if ( valid )
        {
            A.W[p] = W;
            if constexpr ( FRAC )
            {
                Phi = W * vol_i;
                const double ph = 1 - phi_i / vol_H_i;
                A.phi[p] = ph;
                D = ph;
            }
        }
        else
            Phi = 0.0;
    }
    for ( int off = 32; off > 0; off >>= 1 )
    {
        Phi += __shfl_xor( Phi, off );
        D += __shfl_xor( D, off );
    }
    if ( lane == 0 )
    {
        red[0][wv] = Phi;
        red[1][wv] = D;
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        double a = 0.0, d = 0.0;
        for ( int k = 0; k < kSlicesPerBlock; k++ )
        {
            a += red[0][k];
            d += red[1][k];
        }
        partial[2 * t] = a;
        partial[2 * t + 1] = d;
    }
}

template <int MODEL, bool FRAC, bool GEN, int FEAT = 0>
__global__ __launch_bounds__( kBlock ) void kStressRows( const PDArgs* __restrict__ specs, int32_t parity )
{
#pragma clang fp contract( off )
    constexpr bool MT = ( FEAT & kFeatMT ) != 0 && !kIsLps<MODEL>; // LPS: the types are in lpsA / lpsB
    constexpr bool TH = ( FEAT & kFeatTH ) != 0 && !kIsLps<MODEL>;
    constexpr bool EP = ( FEAT & kFeatEP ) != 0 && !kIsLps<MODEL>;
    constexpr int NE = kIsLps<MODEL> ? 2 : 0; // staged lpsA, lpsB
    constexpr int KV = 6 + NE;                // staged vol (GEN)
    constexpr int KT = KV + ( GEN ? 1 : 0 );  // staged temperature (TH)
    constexpr int NC = KT + ( TH ? 1 : 0 );
    constexpr int kW = RowWalk64<NC>::kW;
    __shared__ double win[kSlicesPerBlock][2][NC][kW];
    const PDArgs& A = *specs;
    const int wv = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t slice = xcdRemap( blockIdx.x, gridDim.x ) * kSlicesPerBlock + wv;
    if ( slice >= A.nslices )
        return; // waves are independent: no block-level synchronisation
    const int64_t n = scalarLoad( &A.n );
    const int64_t p0 = slice * kSlice;
    const int64_t p = p0 + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : p0;
    const double* __restrict__ x = scalarLoad( &A.x );
    const double* __restrict__ u = scalarLoad( &A.U[parity] );
    const double* src[NC];
    src[0] = u;
    src[1] = u + n;
    src[2] = u + 2 * n;
    src[3] = x;
    src[4] = x + n;
    src[5] = x + 2 * n;
    if constexpr ( NE )
    {
        src[6] = scalarLoad( &A.lpsA );
        src[7] = scalarLoad( &A.lpsB );
    }
    if constexpr ( GEN )
        src[KV] = scalarLoad( &A.vol );
    if constexpr ( TH )
        src[KT] = scalarLoad( &A.temp );
    const RowWalk64<NC> rw = rowWalk<NC>( A, slice, lane, src );
    const double xi0 = x[pc], xi1 = x[n + pc], xi2 = x[2 * n + pc];
    const double ui0 = u[pc], ui1 = u[n + pc], ui2 = u[2 * n + pc];
    const int64_t row0 = scalarLoad( scalarLoad( &A.slice_ptr ) + slice );
    const uint64_t* __restrict__ ew = scalarLoad( &A.exist ) + row0;
    const uint64_t* __restrict__ iw = scalarLoad( &A.intact ) + row0;
    const uint8_t* __restrict__ mtp = MT ? scalarLoad( &A.mtype ) : nullptr;
    // EPP: the stress pass updates the plastic stretch of every bond it visits
    // (force_models/CabanaPD_PMB.hpp:373-376)
    double* __restrict__ spl = EP ? scalarLoad( &A.splas ) + row0 * kSlice + lane : nullptr;
    const int ti4 = MT ? 4 * static_cast<int>( mtp[pc] ) : 0;
    const double Ti = TH ? A.temp[pc] : 0.0;
    const double c = A.c;
    const double volu = A.vol_value;
    const int infl1 = A.influence;
    double Ai = 0.0, Bi = 0.0;
    if constexpr ( kIsLps<MODEL> )
    {
        Ai = A.lpsA[pc];
        Bi = A.lpsB[pc];
    }
    double s00 = 0.0, s01 = 0.0, s02 = 0.0, s10 = 0.0, s11 = 0.0, s12 = 0.0, s20 = 0.0, s21 = 0.0, s22 = 0.0;
    rw.walk( &win[wv][0][0][0], [&]( int32_t q, const double* __restrict__ wr, int32_t g ) {
        uint64_t live = scalarLoad( ew + q );
        uint64_t intact = ~0ull;
        if constexpr ( FRAC )
        {
            intact = scalarLoad( iw + q );
            if constexpr ( kIsLps<MODEL> )
                live &= intact; // LPS fracture visits unbroken bonds only
        }
        if ( !__builtin_amdgcn_inverse_ballot_w64( live ) )
            return;
        const double xj0 = wr[3 * kW], xj1 = wr[4 * kW], xj2 = wr[5 * kW];
        Bond64 b;
        bond64<MODEL>( xi0, xi1, xi2, ui0, ui1, ui2, xj0, xj1, xj2, wr[0], wr[kW], wr[2 * kW], b );
        const double xi_x = xj0 - xi0;
        const double xi_y = xj1 - xi1;
        const double xi_z = xj2 - xi2;
        const double vj = GEN ? wr[KV * kW] : volu;
        double coeff;
        if constexpr ( kIsLps<MODEL> )
        {
            const double infl = infl1 ? 1.0 / b.xi : 1.0;
            coeff = ( Ai + wr[6 * kW] + b.s * ( Bi + wr[7 * kW] ) ) * infl * b.xi * vj;
            coeff *= 0.5;
        }
        else
        {
            if constexpr ( TH )
                b.s = b.s - ( A.alpha * thermalAvg( A, Ti, wr[KT * kW] ) );
            const double sd = EP ? plasticStretch( spl + q * kSlice, b.s, A.sY ) : b.s;
            coeff = 0.5 * ( ( MT ? A.mat_c[ti4 + mtp[g + lane]] : c ) * sd * vj );
        }
        double fx, fy, fz;
        if constexpr ( FRAC )
        {
            const double mu = static_cast<double>( ( intact >> lane ) & 1ull );
            fx = mu * coeff * b.rx / b.r;
            fy = mu * coeff * b.ry / b.r;
            fz = mu * coeff * b.rz / b.r;
        }
        else
        {
            fx = coeff * b.rx / b.r;
            fy = coeff * b.ry / b.r;
            fz = coeff * b.rz / b.r;
        }
        s00 += fx * xi_x;
        s11 += fy * xi_y;
        s22 += fz * xi_z;
        s01 += fx * xi_y;
        s10 += fy * xi_x;
        s02 += fx * xi_z;
        s20 += fz * xi_x;
        s12 += fy * xi_z;
        s21 += fz * xi_y;
    } );
    if ( !valid )
        return;
    double* __restrict__ st = A.stress;
    st[p] = s00;
    st[n + p] = s01;
    st[2 * n + p] = s02;
    st[3 * n + p] = s10;
    st[4 * n + p] = s11;
    st[5 * n + p] = s12;
    st[6 * n + p] = s20;
    st[7 * n + p] = s21;
    st[8 * n + p] = s22;
}

// ---------------------------------------------------------------------------
// Heat transfer (DynamicTemperature): HeatTransfer::computeHeatTransferFull
// (CabanaPD_HeatTransfer.hpp:41-62, 101-140) -- conduction of particle i over
// its bonds (intact bonds only with fracture), in the bond order of the other
// passes -- then forwardEuler (:66-83) in a second launch, since every
// conduction reads the neighbours' temperatures before the update.  FP64,
// one lane per particle, like the energy pass.
// ---------------------------------------------------------------------------
template <bool FRAC>
__global__ __launch_bounds__( kBlock ) void kHeatConduction( const PDArgs* __restrict__ specs )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int wave = threadIdx.x / kSlice;
    const int64_t slice = blockIdx.x * (int64_t)kSlicesPerBlock + wave;
    const int64_t n = A.n;
    const int64_t p = slice * kSlice + lane;
    if ( slice >= A.nslices || p >= n )
        return;
    const double* __restrict__ x = A.x;
    const double* __restrict__ T = A.temp;
    const double* __restrict__ vol = A.vol;
    const double xi0 = x[p], xi1 = x[n + p], xi2 = x[2 * n + p];
    const double Ti = T[p];
    const int32_t mycnt = A.cnt[p];
    const int64_t base = A.slice_ptr[slice];
    const int32_t qend = slotEnd( A, static_cast<int32_t>( A.slice_ptr[slice + 1] - base ), mycnt );
    const int32_t* __restrict__ colp = A.col + base * kSlice + lane;
    const uint64_t* __restrict__ bits = A.intact + base;
    const double tc = A.thermal_coeff;
    const double th = A.thermal_horizon;
    double conduction = 0.0;
    for ( int32_t q = 0; q < qend; q++ )
    {
        if ( !slotBond( A, base, q, lane, mycnt ) )
            continue;
        if constexpr ( FRAC )
            if ( !( ( bits[q] >> lane ) & 1ull ) )
                continue;
        const int64_t j = colp[q * kSlice];
        const double xi_x = x[j] - xi0;
        const double xi_y = x[n + j] - xi1;
        const double xi_z = x[2 * n + j] - xi2;
        const double xi = sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
        // microconductivity_function (CabanaPD_ForceModels.hpp:302-308)
        const double coeff = A.lin_micro ? 4.0 * tc * ( 1.0 - xi / th ) : tc;
        conduction += coeff * ( T[j] - Ti ) / xi / xi * vol[j];
    }
    A.cond[p] = conduction;
}

// The conduction pass on lattice-row windows (the stencil-slot layout with a
// row table): x and T of the neighbours (and the volumes when they differ)
// staged per row as in the output passes, the bonds in slot order, so the
// sums equal kHeatConduction's bit for bit.
template <bool FRAC, bool GEN>
__global__ __launch_bounds__( kBlock ) void kHeatConductionRows( const PDArgs* __restrict__ specs )
{
#pragma clang fp contract( off )
    constexpr int NC = 4 + ( GEN ? 1 : 0 ); // x, T (, vol)
    constexpr int kW = RowWalk64<NC>::kW;
    __shared__ double win[kSlicesPerBlock][2][NC][kW];
    const PDArgs& A = *specs;
    const int wv = __builtin_amdgcn_readfirstlane( static_cast<int>( threadIdx.x / kSlice ) );
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t slice = xcdRemap( blockIdx.x, gridDim.x ) * kSlicesPerBlock + wv;
    if ( slice >= A.nslices )
        return; // waves are independent: no block-level synchronisation
    const int64_t n = scalarLoad( &A.n );
    const int64_t p0 = slice * kSlice;
    const int64_t p = p0 + lane;
    const bool valid = p < n;
    const int64_t pc = valid ? p : p0;
    const double* __restrict__ x = scalarLoad( &A.x );
    const double* __restrict__ T = scalarLoad( &A.temp );
    const double* src[NC];
    src[0] = x;
    src[1] = x + n;
    src[2] = x + 2 * n;
    src[3] = T;
    if constexpr ( GEN )
        src[4] = scalarLoad( &A.vol );
    const RowWalk64<NC> rw = rowWalk<NC>( A, slice, lane, src );
    const double xi0 = x[pc], xi1 = x[n + pc], xi2 = x[2 * n + pc];
    const double Ti = T[pc];
    const int64_t row0 = scalarLoad( scalarLoad( &A.slice_ptr ) + slice );
    const uint64_t* __restrict__ mw = scalarLoad( FRAC ? &A.intact : &A.exist ) + row0;
    const double tc = A.thermal_coeff;
    const double th = A.thermal_horizon;
    const int lin = A.lin_micro;
    const double volu = A.vol_value;
    double conduction = 0.0;
    rw.walk( &win[wv][0][0][0], [&]( int32_t q, const double* __restrict__ wr, int32_t g ) {
        // intact bonds under fracture (a subset of the existing ones)
        if ( !__builtin_amdgcn_inverse_ballot_w64( scalarLoad( mw + q ) ) )
            return;
        const double xi_x = wr[0] - xi0;
        const double xi_y = wr[kW] - xi1;
        const double xi_z = wr[2 * kW] - xi2;
        const double xi = sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
        // microconductivity_function (CabanaPD_ForceModels.hpp:302-308)
        const double coeff = lin ? 4.0 * tc * ( 1.0 - xi / th ) : tc;
        conduction += coeff * ( wr[3 * kW] - Ti ) / xi / xi * ( GEN ? wr[4 * kW] : volu );
    } );
    if ( valid )
        A.cond[p] = conduction;
}

__global__ __launch_bounds__( 256 ) void kHeatEuler( const PDArgs* __restrict__ specs, double dt )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p >= A.n )
        return;
    A.temp[p] += dt / A.rho[p] / A.cp * A.cond[p];
}

// Velocity-Verlet initial half step on its own (CabanaPD_Integrate.hpp:89-116):
// the first step of a pdk_step call; later steps run it in the force epilogue.
__global__ __launch_bounds__( 256 ) void kInitialHalfStep( const PDArgs* __restrict__ specs,
                                                           int32_t parity )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = A.n;
    if ( p >= n )
        return;
    const double half_dt_m = A.half_dt / A.rho[p];
    const double* u = A.U[parity];
    double* un = A.U[parity ^ 1];
    for ( int d = 0; d < 3; d++ )
    {
        double vv = A.v[d * n + p];
        vv = vv + half_dt_m * A.f[d * n + p];
        A.v[d * n + p] = vv;
        un[d * n + p] = u[d * n + p] + A.dt * vv;
    }
}

// Force BC / body term on the stored f (init at t = 0, CabanaPD_Solver.hpp:272-274).
// ---------------------------------------------------------------------------
// Contact between the particles of a PD specimen: Force<NormalRepulsionModel>::
// computeForceFull (force/CabanaPD_Contact.hpp:44-80) with forceCoeff
// (force_models/CabanaPD_Contact.hpp:82-89), added to the stored PD force in
// list order (computeForce with reset = false, CabanaPD_Solver.hpp:364-366),
// then the force terms and the final half step.  FP64, the oracle's operand
// order; a pair beyond the contact radius adds +0.0 as there.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__( kBlock ) void kPdContact( const PDArgs* __restrict__ specs, PdContactArgs ca,
                                                        int32_t epi, int32_t parity )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t slice = xcdRemap( blockIdx.x, gridDim.x ) * kSlicesPerBlock + threadIdx.x / kSlice;
    if ( slice >= A.nslices )
        return;
    const int64_t n = A.n;
    const int64_t p = slice * kSlice + lane;
    if ( p >= n )
        return; // no cross-lane operations in this kernel
    const double* __restrict__ x = A.x;
    const double* __restrict__ u = A.U[parity];
    const double xi0 = x[p], xi1 = x[n + p], xi2 = x[2 * n + p];
    const double ui0 = u[p], ui1 = u[n + p], ui2 = u[2 * n + p];
    double fx = A.f[p], fy = A.f[n + p], fz = A.f[2 * n + p];
    const int32_t mycnt = ca.cnt[p];
    const int32_t* __restrict__ colp = ca.col + ca.slice_ptr[slice] * kSlice + lane;
    for ( int32_t q = 0; q < mycnt; q++ )
    {
        const int64_t j = colp[q * kSlice];
        Bond64 b;
        getDistance64( xi0, xi1, xi2, ui0, ui1, ui2, x[j], x[n + j], x[2 * n + j], u[j], u[n + j], u[2 * n + j], b );
        double fcx = 0.0, fcy = 0.0, fcz = 0.0;
        if ( b.r < ca.radius )
        {
            const double sc = ( b.r - ca.radius ) / ca.delta;
            const double coeff = 15.0 * ca.c * sc * A.vol[j];
            fcx = coeff * b.rx / b.r;
            fcy = coeff * b.ry / b.r;
            fcz = coeff * b.rz / b.r;
        }
        fx += fcx;
        fy += fcy;
        fz += fcz;
    }
    epilogue( A, p, epi, parity, fx, fy, fz );
}

// Largest |U - u_neigh| (VelocityVerlet<Contact>::initialHalfStep,
// CabanaPD_Integrate.hpp:177-181, with the oracle's sqrt form): block maxima,
// then one block reduces them in block order.
__global__ __launch_bounds__( 256 ) void kMaxDisplacement( const double* __restrict__ U,
                                                           const double* __restrict__ un, int64_t n,
                                                           double* __restrict__ partial )
{
#pragma clang fp contract( off )
    __shared__ double red[256];
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double m = 0.0;
    if ( p < n )
    {
        const double a = U[p] - un[p], b = U[n + p] - un[n + p], c = U[2 * n + p] - un[2 * n + p];
        m = sqrt( a * a + b * b + c * c );
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for ( int w = 128; w > 0; w >>= 1 )
    {
        if ( threadIdx.x < w )
            red[threadIdx.x] = fmax( red[threadIdx.x], red[threadIdx.x + w] );
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        partial[blockIdx.x] = red[0];
}
__global__ __launch_bounds__( 256 ) void kMaxOfPartials( const double* __restrict__ partial, int64_t nb,
                                                         double* __restrict__ out )
{
    __shared__ double red[256];
    double m = 0.0;
    for ( int64_t b = threadIdx.x; b < nb; b += blockDim.x )
        m = fmax( m, partial[b] );
    red[threadIdx.x] = m;
    __syncthreads();
    for ( int w = 128; w > 0; w >>= 1 )
    {
        if ( threadIdx.x < w )
            red[threadIdx.x] = fmax( red[threadIdx.x], red[threadIdx.x + w] );
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        *out = red[0];
}

__global__ __launch_bounds__( 256 ) void kApplyForceTerms( const PDArgs* __restrict__ specs )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = A.n;
    if ( p >= n )
        return;
    for ( int d = 0; d < 3; d++ )
    {
        double fv = A.f[d * n + p];
        if ( A.body )
            fv = fv + A.body[d * n + p];
        if ( A.bc_kind >= 0 && A.bcflag[p] )
            fv = ( A.bc_kind == PDK_BC_FORCE_VALUE ) ? A.bc_value : fv + A.bc_value;
        A.f[d * n + p] = fv;
    }
}

// Prescribed values (boundary conditions and body terms with force_update ==
// false, CabanaPD_Solver.hpp:255-257, 350-352): member m (device index
// idx[m]) gets a[m][c] + b[m][c] t in component c of dst ([ncomp][n]).
__global__ __launch_bounds__( 256 ) void kPrescribe( const int64_t* __restrict__ idx, int64_t count,
                                                     const double* __restrict__ a, const double* __restrict__ b,
                                                     int32_t ncomp, double t, double* __restrict__ dst, int64_t n )
{
#pragma clang fp contract( off )
    const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( m >= count )
        return;
    const int64_t p = idx[m];
    for ( int32_t c = 0; c < ncomp; c++ )
    {
        const int64_t e = m * ncomp + c;
        dst[c * n + p] = b ? a[e] + b[e] * t : a[e];
    }
}

using ForceFn = void ( * )( const PDArgs*, const int32_t*, const int64_t*, int32_t, int32_t, int64_t, int32_t );

template <int MODEL, bool FRAC, int PREC, bool SLOTS, bool MM = false>
ForceFn pickUnroll( int unroll )
{
    switch ( unroll )
    {
    case 1: return kForce<MODEL, FRAC, PREC, 1, SLOTS, MM>;
    case 2: return kForce<MODEL, FRAC, PREC, 2, SLOTS, MM>;
    case 8: return kForce<MODEL, FRAC, PREC, 8, SLOTS, MM>;
    default: return kForce<MODEL, FRAC, PREC, 4, SLOTS, MM>;
    }
}

template <int MODEL, bool SLOTS>
ForceFn pickForceModel( const PDArgs& a )
{
    const bool mixed = a.precision == PDK_MIXED;
    if constexpr ( kIsLin<MODEL> ) // no fracture form
        return mixed ? pickUnroll<MODEL, false, kMixed, SLOTS>( a.unroll )
                     : pickUnroll<MODEL, false, kF64, SLOTS>( a.unroll );
    else if ( mixed && a.mtype )
        return a.fracture ? pickUnroll<MODEL, true, kMixed, SLOTS, true>( a.unroll )
                          : pickUnroll<MODEL, false, kMixed, SLOTS, true>( a.unroll );
    else if ( a.fracture )
        return mixed ? pickUnroll<MODEL, true, kMixed, SLOTS>( a.unroll )
                     : pickUnroll<MODEL, true, kF64, SLOTS>( a.unroll );
    else
        return mixed ? pickUnroll<MODEL, false, kMixed, SLOTS>( a.unroll )
                     : pickUnroll<MODEL, false, kF64, SLOTS>( a.unroll );
}

template <bool SLOTS>
ForceFn pickForceForm( const PDArgs& a )
{
    switch ( a.model )
    {
    case kPMB: return pickForceModel<kPMB, SLOTS>( a );
    case kLPS: return pickForceModel<kLPS, SLOTS>( a );
    case kLinPMB: return pickForceModel<kLinPMB, SLOTS>( a );
    default: return pickForceModel<kLinLPS, SLOTS>( a );
    }
}

template <int MODEL, bool GEN>
ForceFn pickRowsModel( const PDArgs& a )
{
    const bool mixed = a.precision == PDK_MIXED;
    if constexpr ( kIsLin<MODEL> ) // no fracture form
        return mixed ? kForceRows<MODEL, false, kMixed, GEN> : kForceRows<MODEL, false, kF64, GEN>;
    else if ( a.fracture )
        return mixed ? kForceRows<MODEL, true, kMixed, GEN> : kForceRows<MODEL, true, kF64, GEN>;
    else
        return mixed ? kForceRows<MODEL, false, kMixed, GEN> : kForceRows<MODEL, false, kF64, GEN>;
}

template <bool GEN>
ForceFn pickForceRowsGen( const PDArgs& a )
{
    switch ( a.model )
    {
    case kPMB: return pickRowsModel<kPMB, GEN>( a );
    case kLPS: return pickRowsModel<kLPS, GEN>( a );
    case kLinPMB: return pickRowsModel<kLinPMB, GEN>( a );
    default: return pickRowsModel<kLinLPS, GEN>( a );
    }
}

// The general variant stages per-particle volumes and no-fail flags; the
// other one uses the uniform volume and no no-fail particles (exact
// shortcuts detected at upload).
bool rowsGeneral( const PDArgs& a )
{
    return !a.vol_uniform || ( a.fracture && a.any_nofail );
}

ForceFn pickForceRows( const PDArgs& a )
{
    return rowsGeneral( a ) ? pickForceRowsGen<true>( a ) : pickForceRowsGen<false>( a );
}

// Per-particle features of the force pass (kFeat* bits): several materials,
// temperature, plastic stretch (the LPS model has the first only).
int rowFeatures( const PDArgs& a )
{
    const bool lps = a.model == kLPS || a.model == kLinLPS;
    return ( a.mtype ? kFeatMT : 0 ) | ( !lps && a.temp ? kFeatTH : 0 ) | ( !lps && a.splas ? kFeatEP : 0 );
}

template <int MODEL, int FEAT>
ForceFn pickRowsFeat( const PDArgs& a )
{
    const bool mixed = a.precision == PDK_MIXED;
    if constexpr ( ( FEAT & kFeatEP ) != 0 ) // ElasticPerfectlyPlastic exists with fracture only
        return mixed ? kForceRowsF<MODEL, true, kMixed, FEAT> : kForceRowsF<MODEL, true, kF64, FEAT>;
    else if ( a.fracture )
        return mixed ? kForceRowsF<MODEL, true, kMixed, FEAT> : kForceRowsF<MODEL, true, kF64, FEAT>;
    else
        return mixed ? kForceRowsF<MODEL, false, kMixed, FEAT> : kForceRowsF<MODEL, false, kF64, FEAT>;
}

// The combinations the feature instances cover: several materials (PMB or
// LPS), or a temperature and / or a plastic stretch (PMB); the others take
// the gathers.
ForceFn pickForceRowsFeat( const PDArgs& a )
{
    switch ( rowFeatures( a ) )
    {
    case kFeatMT: return a.model == kLPS ? pickRowsFeat<kLPS, kFeatMT>( a ) : pickRowsFeat<kPMB, kFeatMT>( a );
    case kFeatTH: return pickRowsFeat<kPMB, kFeatTH>( a );
    case kFeatEP: return pickRowsFeat<kPMB, kFeatEP>( a );
    default: return pickRowsFeat<kPMB, kFeatTH | kFeatEP>( a );
    }
}

ForceFn pickForce( const PDArgs& a )
{
    if ( useRowWindows( a ) )
        return rowFeatures( a ) ? pickForceRowsFeat( a ) : pickForceRows( a );
    return a.exist ? pickForceForm<true>( a ) : pickForceForm<false>( a );
}

thread_local char g_force_name[64];
thread_local char g_lps_name[64];
thread_local char g_energy_name[64];
thread_local char g_stress_name[64];

} // namespace

int normalizeUnroll( int u )
{
    return ( u == 1 || u == 2 || u == 8 ) ? u : 4;
}

bool rowFeaturesSupported( int32_t model, bool pair_tables, bool thermal, bool plastic )
{
    if ( pair_tables )
        return ( model == kPMB || model == kLPS ) && !thermal && !plastic;
    return model == kPMB || !( thermal || plastic );
}

const char* forceKernelName( const PDArgs& a )
{
    if ( useRowWindows( a ) && rowFeatures( a ) )
    {
        std::snprintf( g_force_name, sizeof( g_force_name ), "kForceRowsF<%d, %s, %d, %d>", a.model,
                       a.fracture ? "true" : "false", a.precision == PDK_MIXED ? 1 : 0, rowFeatures( a ) );
        return g_force_name;
    }
    if ( useRowWindows( a ) )
    {
        std::snprintf( g_force_name, sizeof( g_force_name ), "kForceRows<%d, %s, %d, %s>", a.model,
                       a.fracture ? "true" : "false", a.precision == PDK_MIXED ? 1 : 0,
                       rowsGeneral( a ) ? "true" : "false" );
        return g_force_name;
    }
    const bool mm = a.precision == PDK_MIXED && a.mtype != nullptr && ( a.model == kPMB || a.model == kLPS );
    std::snprintf( g_force_name, sizeof( g_force_name ), "kForce<%d, %s, %d, %d, %s%s>", a.model,
                   a.fracture ? "true" : "false", a.precision == PDK_MIXED ? 1 : 0, normalizeUnroll( a.unroll ),
                   a.exist ? "true" : "false", mm ? ", true" : "" );
    return g_force_name;
}
const char* lpsKernelName( const PDArgs& a )
{
    if ( a.rows != nullptr && a.nrows > 0 )
    {
        std::snprintf( g_lps_name, sizeof( g_lps_name ), "kLpsStateRows<%s, %d, %s>",
                       a.fracture ? "true" : "false", a.precision == PDK_MIXED ? 1 : 0,
                       a.vol_uniform ? "false" : "true" );
        return g_lps_name;
    }
    return a.fracture ? ( a.precision == PDK_MIXED ? "kLpsState<true, 1>" : "kLpsState<true, 0>" )
                      : ( a.precision == PDK_MIXED ? "kLpsState<false, 1>" : "kLpsState<false, 0>" );
}
// The output passes take the row windows wherever the layout has a row table
// and the specimen has no per-particle material type, temperature or plastic
// stretch (the same results either way).
static bool outputRows( const PDArgs& a )
{
    if ( a.rows == nullptr || a.nrows <= 0 || a.exist == nullptr )
        return false;
    const bool lps = a.model == kLPS || a.model == kLinLPS;
    const int feat = ( a.mtype ? kFeatMT : 0 ) | ( !lps && a.temp ? kFeatTH : 0 ) | ( !lps && a.splas ? kFeatEP : 0 );
    switch ( feat )
    {
    case 0: return true;
    case kFeatMT: return a.model == kPMB || a.model == kLPS;
    default: return ( feat & kFeatMT ) == 0 && a.model == kPMB; // temperature and / or plastic stretch
    }
}
static int outputFeatures( const PDArgs& a )
{
    const bool lps = a.model == kLPS || a.model == kLinLPS;
    return ( a.mtype ? kFeatMT : 0 ) | ( !lps && a.temp ? kFeatTH : 0 ) | ( !lps && a.splas ? kFeatEP : 0 );
}

const char* energyKernelName( const PDArgs& a )
{
    if ( outputRows( a ) )
    {
        const int feat = outputFeatures( a );
        if ( feat )
            std::snprintf( g_energy_name, sizeof( g_energy_name ), "kEnergyRows<%d, %s, true, %d>", a.model,
                           a.fracture ? "true" : "false", feat );
        else
            std::snprintf( g_energy_name, sizeof( g_energy_name ), "kEnergyRows<%d, %s, %s>", a.model,
                           a.fracture && !( a.model == kLinPMB || a.model == kLinLPS ) ? "true" : "false",
                           a.vol_uniform ? "false" : "true" );
        return g_energy_name;
    }
    if ( a.model == kLinPMB )
        return "kEnergy<4, false>";
    if ( a.model == kLinLPS )
        return "kEnergy<5, false>";
    return a.model == kLPS ? ( a.fracture ? "kEnergy<1, true>" : "kEnergy<1, false>" )
                           : ( a.fracture ? "kEnergy<0, true>" : "kEnergy<0, false>" );
}

static unsigned sliceBlocks( const PDArgs& a )
{
    return static_cast<unsigned>( ( a.nslices + kSlicesPerBlock - 1 ) / kSlicesPerBlock );
}

void launchForce( const PDArgs* args_dev, const PDArgs& a, int32_t epilogue_mode, int32_t parity,
                  hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    const int32_t accumulate = epilogue_mode < 0 ? 1 : 0;
    const int32_t epi = epilogue_mode < 0 ? kEpiStoreOnly : epilogue_mode;
    ForceFn fn = pickForce( a );
    hipLaunchKernelGGL( fn, dim3( sliceBlocks( a ) ), dim3( kBlock ), 0, st, args_dev, nullptr, nullptr, epi,
                        parity, int64_t( 0 ), accumulate );
    PDK_HIP( hipGetLastError() );
}

using LpsFn = void ( * )( const PDArgs*, const int32_t*, const int64_t*, int32_t, int64_t, int32_t, int32_t );

template <bool GEN>
LpsFn pickLpsRows( const PDArgs& a )
{
    return a.fracture ? ( a.precision == PDK_MIXED ? kLpsStateRows<true, kMixed, GEN> : kLpsStateRows<true, kF64, GEN> )
                      : ( a.precision == PDK_MIXED ? kLpsStateRows<false, kMixed, GEN>
                                                   : kLpsStateRows<false, kF64, GEN> );
}

// The LPS state pass follows the layout (row windows wherever a row table
// exists), not the force-pass tuning knob; both kernels give the same results.
LpsFn pickLps( const PDArgs& a )
{
    if ( a.rows != nullptr && a.nrows > 0 )
        return a.vol_uniform ? pickLpsRows<false>( a ) : pickLpsRows<true>( a );
    return a.fracture ? ( a.precision == PDK_MIXED ? kLpsState<true, kMixed> : kLpsState<true, kF64> )
                      : ( a.precision == PDK_MIXED ? kLpsState<false, kMixed> : kLpsState<false, kF64> );
}

void launchLpsState( const PDArgs* args_dev, const PDArgs& a, bool weighted_volume, bool dilatation,
                     int32_t parity, hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    LpsFn fn = pickLps( a );
    hipLaunchKernelGGL( fn, dim3( sliceBlocks( a ) ), dim3( kBlock ), 0, st, args_dev, nullptr, nullptr, parity,
                        int64_t( 0 ), int32_t( weighted_volume ), int32_t( dilatation ) );
    PDK_HIP( hipGetLastError() );
}

// The packed steps of a replayed graph done: the device launch counter
// advances by their number (the graph's last node).
__global__ void kAdvanceLaunch( int64_t* __restrict__ ctr, int64_t steps )
{
    if ( blockIdx.x == 0 && threadIdx.x == 0 )
        *ctr = *ctr + steps;
}
// The device launch counter set to L before replays.
__global__ void kSetLaunch( int64_t* __restrict__ ctr, int64_t L )
{
    if ( blockIdx.x == 0 && threadIdx.x == 0 )
        *ctr = L;
}

void launchAdvanceLaunch( int64_t* ctr, int64_t steps, hipStream_t st )
{
    hipLaunchKernelGGL( kAdvanceLaunch, dim3( 1 ), dim3( 64 ), 0, st, ctr, steps );
    PDK_HIP( hipGetLastError() );
}

void launchSetLaunch( int64_t* ctr, int64_t L, hipStream_t st )
{
    hipLaunchKernelGGL( kSetLaunch, dim3( 1 ), dim3( 64 ), 0, st, ctr, L );
    PDK_HIP( hipGetLastError() );
}

void launchForceBatched( const PDArgs* specs_dev, const PDArgs& variant, const int32_t* tile_spec,
                         const int64_t* tile_slice, int64_t ntiles, int64_t launch_index, hipStream_t st )
{
    if ( ntiles == 0 )
        return;
    ForceFn fn = pickForce( variant );
    hipLaunchKernelGGL( fn, dim3( static_cast<unsigned>( ntiles ) ), dim3( kBlock ), 0, st, specs_dev, tile_spec,
                        tile_slice, int32_t( kEpiFinalInitial ), static_cast<int32_t>( launch_index & 1 ),
                        launch_index, int32_t( 0 ) );
    PDK_HIP( hipGetLastError() );
}

void launchLpsStateBatched( const PDArgs* specs_dev, const PDArgs& variant, const int32_t* tile_spec,
                            const int64_t* tile_slice, int64_t ntiles, int64_t launch_index, hipStream_t st )
{
    if ( ntiles == 0 )
        return;
    LpsFn fn = pickLps( variant );
    hipLaunchKernelGGL( fn, dim3( static_cast<unsigned>( ntiles ) ), dim3( kBlock ), 0, st, specs_dev, tile_spec,
                        tile_slice, static_cast<int32_t>( launch_index & 1 ), launch_index,
                        int32_t( variant.fracture ? 1 : 0 ), int32_t( 1 ) );
    PDK_HIP( hipGetLastError() );
}

using EnergyFn = void ( * )( const PDArgs*, int32_t, const double*, const double*, double* );
template <bool GEN>
EnergyFn pickEnergyRows( const PDArgs& a )
{
    switch ( a.model )
    {
    case kLinPMB: return kEnergyRows<kLinPMB, false, GEN>;
    case kLinLPS: return kEnergyRows<kLinLPS, false, GEN>;
    case kLPS: return a.fracture ? kEnergyRows<kLPS, true, GEN> : kEnergyRows<kLPS, false, GEN>;
    default: return a.fracture ? kEnergyRows<kPMB, true, GEN> : kEnergyRows<kPMB, false, GEN>;
    }
}
using StressFn = void ( * )( const PDArgs*, int32_t );
template <bool GEN>
StressFn pickStressRows( const PDArgs& a )
{
    switch ( a.model )
    {
    case kLinPMB: return kStressRows<kLinPMB, false, GEN>;
    case kLinLPS: return kStressRows<kLinLPS, false, GEN>;
    case kLPS: return a.fracture ? kStressRows<kLPS, true, GEN> : kStressRows<kLPS, false, GEN>;
    default: return a.fracture ? kStressRows<kPMB, true, GEN> : kStressRows<kPMB, false, GEN>;
    }
}

// Feature specimens (the combinations outputRows admits) stage the volumes
// always (GEN), which gives the same values when they are uniform.
EnergyFn pickEnergyRowsFeat( const PDArgs& a )
{
    switch ( outputFeatures( a ) )
    {
    case kFeatMT:
        if ( a.model == kLPS )
            return a.fracture ? kEnergyRows<kLPS, true, true, kFeatMT> : kEnergyRows<kLPS, false, true, kFeatMT>;
        return a.fracture ? kEnergyRows<kPMB, true, true, kFeatMT> : kEnergyRows<kPMB, false, true, kFeatMT>;
    case kFeatTH:
        return a.fracture ? kEnergyRows<kPMB, true, true, kFeatTH> : kEnergyRows<kPMB, false, true, kFeatTH>;
    case kFeatEP: return kEnergyRows<kPMB, true, true, kFeatEP>;
    default: return kEnergyRows<kPMB, true, true, kFeatTH | kFeatEP>;
    }
}
StressFn pickStressRowsFeat( const PDArgs& a )
{
    switch ( outputFeatures( a ) )
    {
    case kFeatMT:
        if ( a.model == kLPS )
            return a.fracture ? kStressRows<kLPS, true, true, kFeatMT> : kStressRows<kLPS, false, true, kFeatMT>;
        return a.fracture ? kStressRows<kPMB, true, true, kFeatMT> : kStressRows<kPMB, false, true, kFeatMT>;
    case kFeatTH:
        return a.fracture ? kStressRows<kPMB, true, true, kFeatTH> : kStressRows<kPMB, false, true, kFeatTH>;
    case kFeatEP: return kStressRows<kPMB, true, true, kFeatEP>;
    default: return kStressRows<kPMB, true, true, kFeatTH | kFeatEP>;
    }
}

int64_t energyBlocks( const PDArgs& a )
{
    return sliceBlocks( a );
}

void launchEnergy( const PDArgs* args_dev, const PDArgs& a, int32_t parity, double* partial_dev,
                   int64_t nblocks, hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    auto fn = a.model == kLinPMB ? kEnergy<kLinPMB, false>
              : a.model == kLinLPS ? kEnergy<kLinLPS, false>
              : a.model == kLPS    ? ( a.fracture ? kEnergy<kLPS, true> : kEnergy<kLPS, false> )
                                   : ( a.fracture ? kEnergy<kPMB, true> : kEnergy<kPMB, false> );
    if ( outputRows( a ) )
        fn = outputFeatures( a ) ? pickEnergyRowsFeat( a )
                                 : ( a.vol_uniform ? pickEnergyRows<false>( a ) : pickEnergyRows<true>( a ) );
    hipLaunchKernelGGL( fn, dim3( static_cast<unsigned>( nblocks ) ), dim3( kBlock ), 0, st, args_dev, parity,
                        a.m, a.theta, partial_dev );
    PDK_HIP( hipGetLastError() );
}

void launchHeatTransfer( const PDArgs* args_dev, const PDArgs& a, double dt, hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    auto fn = a.fracture ? kHeatConduction<true> : kHeatConduction<false>;
    if ( a.rows != nullptr && a.nrows > 0 && a.exist != nullptr )
        fn = a.vol_uniform ? ( a.fracture ? kHeatConductionRows<true, false> : kHeatConductionRows<false, false> )
                           : ( a.fracture ? kHeatConductionRows<true, true> : kHeatConductionRows<false, true> );
    hipLaunchKernelGGL( fn, dim3( sliceBlocks( a ) ), dim3( kBlock ), 0, st, args_dev );
    PDK_HIP( hipGetLastError() );
    hipLaunchKernelGGL( kHeatEuler, dim3( static_cast<unsigned>( ( a.n + 255 ) / 256 ) ), dim3( 256 ), 0, st, args_dev,
                        dt );
    PDK_HIP( hipGetLastError() );
}

void launchStress( const PDArgs* args_dev, const PDArgs& a, int32_t parity, hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    auto fn = a.model == kLinPMB ? kStress<kLinPMB, false>
              : a.model == kLinLPS ? kStress<kLinLPS, false>
              : a.model == kLPS    ? ( a.fracture ? kStress<kLPS, true> : kStress<kLPS, false> )
                                   : ( a.fracture ? kStress<kPMB, true> : kStress<kPMB, false> );
    if ( outputRows( a ) )
        fn = outputFeatures( a ) ? pickStressRowsFeat( a )
                                 : ( a.vol_uniform ? pickStressRows<false>( a ) : pickStressRows<true>( a ) );
    hipLaunchKernelGGL( fn, dim3( sliceBlocks( a ) ), dim3( kBlock ), 0, st, args_dev, parity );
    PDK_HIP( hipGetLastError() );
}

namespace
{
// Per stencil entry (FP32): 1/|xi|, |xi|^2 and the LPS influence products
// infl |xi| and infl |xi|^2 (infl = 1/|xi| for influence 1, else 1).
__global__ void kStencilAux( const float4* __restrict__ table, int32_t S, int32_t influence, float4* __restrict__ aux )
{
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if ( q < S )
    {
        const float w = table[q].w;
        const float xinv = __builtin_amdgcn_rcpf( w );
        aux[q] = make_float4( xinv, w * w, influence ? 1.0f : w, influence ? w : w * w );
    }
}
__global__ void kStencilRec( const float4* __restrict__ table, const float4* __restrict__ aux,
                             const double* __restrict__ xiv, int32_t S, double* __restrict__ rec )
{
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if ( q < S )
    {
        double* r = rec + kRecStride * q;
        r[kRecXiX] = xiv[4 * q + kXiX];
        r[kRecXi2] = xiv[4 * q + kXi2];
        r[kRecXiY] = xiv[4 * q + kXiY];
        r[kRecXiZ] = xiv[4 * q + kXiZ];
        float* f = reinterpret_cast<float*>( r + kRecF );
        const float4 a = aux[q];
        f[0] = table[q].w;
        f[1] = a.z;
        f[2] = a.w;
        f[3] = a.x;
    }
}
} // namespace
void launchStencilRec( const float4* table, const float4* aux, const double* xiv, int32_t S, double* rec,
                       hipStream_t st )
{
    if ( S <= 0 )
        return;
    hipLaunchKernelGGL( kStencilRec, dim3( ( S + 255 ) / 256 ), dim3( 256 ), 0, st, table, aux, xiv, S, rec );
    PDK_HIP( hipGetLastError() );
}
void launchStencilAux( const float4* table, int32_t S, int32_t influence, float4* aux, hipStream_t st )
{
    if ( S <= 0 )
        return;
    hipLaunchKernelGGL( kStencilAux, dim3( ( S + 255 ) / 256 ), dim3( 256 ), 0, st, table, S, influence, aux );
    PDK_HIP( hipGetLastError() );
}

const char* stressKernelName( const PDArgs& a )
{
    if ( outputRows( a ) )
    {
        const int feat = outputFeatures( a );
        if ( feat )
            std::snprintf( g_stress_name, sizeof( g_stress_name ), "kStressRows<%d, %s, true, %d>", a.model,
                           a.fracture ? "true" : "false", feat );
        else
            std::snprintf( g_stress_name, sizeof( g_stress_name ), "kStressRows<%d, %s, %s>", a.model,
                           a.fracture && !( a.model == kLinPMB || a.model == kLinLPS ) ? "true" : "false",
                           a.vol_uniform ? "false" : "true" );
        return g_stress_name;
    }
    if ( a.model == kLinPMB )
        return "kStress<4, false>";
    if ( a.model == kLinLPS )
        return "kStress<5, false>";
    return a.model == kLPS ? ( a.fracture ? "kStress<1, true>" : "kStress<1, false>" )
                           : ( a.fracture ? "kStress<0, true>" : "kStress<0, false>" );
}

void launchInitialHalfStep( const PDArgs* args_dev, const PDArgs& a, int32_t parity, hipStream_t st )
{
    if ( a.n == 0 )
        return;
    hipLaunchKernelGGL( kInitialHalfStep, dim3( static_cast<unsigned>( ( a.n + 255 ) / 256 ) ), dim3( 256 ), 0, st,
                        args_dev, parity );
    PDK_HIP( hipGetLastError() );
}

void launchPdContact( const PDArgs* args_dev, const PDArgs& a, const PdContactArgs& ca, int32_t epi, int32_t parity,
                      hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    hipLaunchKernelGGL( kPdContact, dim3( sliceBlocks( a ) ), dim3( kBlock ), 0, st, args_dev, ca, epi, parity );
    PDK_HIP( hipGetLastError() );
}

int64_t maxDisplacementBlocks( int64_t n )
{
    return ( n + 255 ) / 256;
}

void launchMaxDisplacement( const double* U, const double* u_neigh, int64_t n, double* partial, double* out_dev,
                            hipStream_t st )
{
    const int64_t nb = maxDisplacementBlocks( n );
    if ( nb == 0 )
        return;
    hipLaunchKernelGGL( kMaxDisplacement, dim3( static_cast<unsigned>( nb ) ), dim3( 256 ), 0, st, U, u_neigh, n,
                        partial );
    PDK_HIP( hipGetLastError() );
    hipLaunchKernelGGL( kMaxOfPartials, dim3( 1 ), dim3( 256 ), 0, st, partial, nb, out_dev );
    PDK_HIP( hipGetLastError() );
}

void launchApplyForceTerms( const PDArgs* args_dev, const PDArgs& a, hipStream_t st )
{
    if ( a.n == 0 )
        return;
    hipLaunchKernelGGL( kApplyForceTerms, dim3( static_cast<unsigned>( ( a.n + 255 ) / 256 ) ), dim3( 256 ), 0,
                        st, args_dev );
    PDK_HIP( hipGetLastError() );
}

void launchPrescribed( const int64_t* idx_dev, int64_t count, const double* a_dev, const double* b_dev, int32_t ncomp,
                       double t, double* dst_dev, int64_t n, hipStream_t st )
{
    if ( count == 0 )
        return;
    hipLaunchKernelGGL( kPrescribe, dim3( static_cast<unsigned>( ( count + 255 ) / 256 ) ), dim3( 256 ), 0, st,
                        idx_dev, count, a_dev, b_dev, ncomp, t, dst_dev, n );
    PDK_HIP( hipGetLastError() );
}

} // namespace pdk
