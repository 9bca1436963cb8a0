// pdk_contact.hip -- DEM contact step: Hertz and Hertz-JKR pair forces over a
// cell-binned contact list on current positions.
//
// Reference: Force<HertzianModel> (force/CabanaPD_Contact.hpp:109-172) with
// HertzianModel (force_models/CabanaPD_Hertzian.hpp:58-87) or
// HertzianJKRModel (force_models/CabanaPD_HertzianJKR.hpp:82-176), the list
// rebuild of Neighbor::update (CabanaPD_Neighbor.hpp:106-132) and
// VelocityVerlet<Contact> (CabanaPD_Integrate.hpp:147-193).
//
// Extension beyond the reference (which has one radius and equal masses):
// with per-particle radii the pair uses R* = R_i R_j / (R_i + R_j), overlap
// R_i + R_j - r and effective mass m_i m_j / (m_i + m_j).  Uniform specs use
// the reference's expressions unchanged.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pdk_serving.hpp"

namespace pdk
{
namespace
{
constexpr double kPi = 3.141592653589793238462643383279502884;

struct PairK
{
    double Rs, coeff_h_n, delta_tear, two_r, ms;
};

__device__ double jkrPatchRadius( const PDArgs& A, double Rs, double delta_n )
{
#pragma clang fp contract( off )
    // force_models/CabanaPD_HertzianJKR.hpp:141-176
    const double nu = A.nu, E = A.E, Gamma = A.Gamma;
    double c0 = pow( Rs * delta_n, 2.0 );
    double c1 = -4.0 * ( 1.0 - pow( nu, 2.0 ) ) * kPi * Gamma * pow( Rs, 2.0 ) / E;
    double c2 = -2.0 * Rs * delta_n;
    double P = -pow( c2, 2.0 ) / 12.0 - c0;
    double Q = -pow( c2, 3.0 ) / 108.0 + c0 * c2 / 3.0 - pow( c1, 2.0 ) / 8.0;
    double U = pow( -Q / 2.0 + pow( fmax( 0.0, pow( Q, 2.0 ) / 4.0 + pow( P, 3.0 ) / 27.0 ), 1.0 / 2.0 ), 1.0 / 3.0 );
    double sv = 0.0;
    if ( P != 0.0 )
        sv = -5.0 / 6.0 * c2 + U - P / ( 3.0 * U );
    else
        sv = -5.0 / 6.0 * c2 - pow( Q, 1.0 / 3.0 );
    double w = pow( fmax( 0.0, c2 + 2.0 * sv ), 1.0 / 2.0 );
    double lambda = c1 / ( 2.0 * w + 1e-14 );
    return 1.0 / 2.0 * ( w + pow( fmax( 0.0, pow( w, 2.0 ) - 4.0 * ( c2 + sv + lambda ) ), 1.0 / 2.0 ) );
}

template <bool JKR>
__device__ double normalForce( const PDArgs& A, const PairK& pk, double r, double vn, double vol )
{
#pragma clang fp contract( off )
    if constexpr ( !JKR )
    {
        const double delta_n = ( r - pk.two_r );
        double coeff = 0.0;
        if ( delta_n < 0.0 )
            coeff = fmin( 0.0, -pk.coeff_h_n * pow( fabs( delta_n ), 3.0 / 2.0 ) );
        coeff /= vol;
        double Sn = 0.0;
        if ( delta_n < 0.0 )
            Sn = 2.0 * A.Es * sqrt( pk.Rs * fabs( delta_n ) );
        coeff -= A.coeff_h_d * sqrt( Sn * pk.ms ) * vn / vol;
        return coeff;
    }
    else
    {
        const double delta_n = ( pk.two_r - r );
        double coeff = 0.0;
        if ( delta_n >= pk.delta_tear )
        {
            double a = jkrPatchRadius( A, pk.Rs, delta_n );
            double a3 = pow( a, 3.0 );
            coeff = -1.0 * ( pk.coeff_h_n * a3 - pow( 8.0 * A.Gamma * kPi * A.Es * a3, 1.0 / 2.0 ) );
            double Sn = 2.0 * A.Es * a;
            coeff -= A.coeff_h_d * sqrt( Sn * pk.ms ) * vn;
            if ( delta_n <= 0.0 && vn <= 0.0 )
                coeff = 0.0;
        }
        coeff /= vol;
        return coeff;
    }
}

template <bool POLY, bool JKR>
__global__ __launch_bounds__( kBlock ) void kContact( const PDArgs* __restrict__ specs, int32_t epi,
                                                      int32_t parity )
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
    const double* __restrict__ v = A.v;
    const double xi0 = x[p], xi1 = x[n + p], xi2 = x[2 * n + p];
    const double ui0 = u[p], ui1 = u[n + p], ui2 = u[2 * n + p];
    const double vi0 = v[p], vi1 = v[n + p], vi2 = v[2 * n + p];
    const double vol_i = A.vol[p];
    const double rho_i = A.rho[p];
    const double Ri = POLY ? A.radius[p] : 0.0;
    const int32_t mycnt = A.cnt[p];
    const int64_t base = A.slice_ptr[slice];
    const int32_t* __restrict__ colp = A.col + base * kSlice + lane;
    PairK pk;
    pk.Rs = A.Rs;
    pk.coeff_h_n = A.coeff_h_n;
    pk.delta_tear = A.delta_tear;
    pk.two_r = A.two_r;
    pk.ms = ( rho_i * vol_i ) / 2.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for ( int32_t q = 0; q < mycnt; q++ )
    {
        const int64_t j = colp[q * kSlice];
        const double rx = ( x[j] - xi0 ) + ( u[j] - ui0 );
        const double ry = ( x[n + j] - xi1 ) + ( u[n + j] - ui1 );
        const double rz = ( x[2 * n + j] - xi2 ) + ( u[2 * n + j] - ui2 );
        const double r = sqrt( rx * rx + ry * ry + rz * rz );
        const double nx = rx / r, ny = ry / r, nz = rz / r;
        const double vx = v[j] - vi0, vy = v[n + j] - vi1, vz = v[2 * n + j] - vi2;
        const double vn = vx * nx + vy * ny + vz * nz;
        if constexpr ( POLY )
        {
            const double Rj = A.radius[j];
            pk.Rs = Ri * Rj / ( Ri + Rj );
            pk.two_r = Ri + Rj;
            if constexpr ( JKR )
            {
                pk.coeff_h_n = 4.0 / 3.0 * A.Es / pk.Rs;
                const double a0 = pow( 9.0 / 2.0 * A.Gamma * kPi / A.Es * pow( pk.Rs, 2.0 ), 1.0 / 3.0 );
                pk.delta_tear = -1.0 / 2.0 * pow( 6.0, -1.0 / 3.0 ) * pow( a0, 2.0 ) / pk.Rs;
            }
            else
                pk.coeff_h_n = 4.0 / 3.0 * A.Es * sqrt( pk.Rs );
            const double mi = rho_i * vol_i, mj = A.rho[j] * A.vol[j];
            pk.ms = mi * mj / ( mi + mj );
        }
        const double fn = normalForce<JKR>( A, pk, r, vn, vol_i );
        fx += fn * nx;
        fy += fn * ny;
        fz += fn * nz;
        double ftx = 0.0, fty = 0.0, ftz = 0.0;
        if constexpr ( JKR )
        {
            // force_models/CabanaPD_HertzianJKR.hpp:114-129
            // Without friction or normal force every term is a signed zero,
            // which leaves the sum unchanged: skip the three tanh.
            const double fna = fabs( fn );
            const double mu = A.mu_f;
            if ( mu != 0.0 && fna != 0.0 )
            {
                const double vtx = vx - vn * nx, vty = vy - vn * ny, vtz = vz - vn * nz;
                const double vt = sqrt( vtx * vtx + vty * vty + vtz * vtz );
                ftx = ( vt > 0 ) ? fabs( mu * tanh( 10 * vtx ) ) * fna * vtx / vt : 0.0;
                fty = ( vt > 0 ) ? fabs( mu * tanh( 10 * vty ) ) * fna * vty / vt : 0.0;
                ftz = ( vt > 0 ) ? fabs( mu * tanh( 10 * vtz ) ) * fna * vtz / vt : 0.0;
            }
        }
        fx += ftx;
        fy += fty;
        fz += ftz;
    }
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
                fx = fy = fz = A.bc_value;
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
    // final half step (CabanaPD_Integrate.hpp:128-134)
    const double half_dt_m = A.half_dt / rho_i;
    A.v[p] = vi0 + half_dt_m * fx;
    A.v[n + p] = vi1 + half_dt_m * fy;
    A.v[2 * n + p] = vi2 + half_dt_m * fz;
}

// VelocityVerlet<Contact>::initialHalfStep (CabanaPD_Integrate.hpp:154-193):
// the half kick, the drift into the other displacement buffer, and the block
// maxima of |u - u_build|.
__global__ __launch_bounds__( 256 ) void kContactInitial( const PDArgs* __restrict__ specs, int32_t parity,
                                                          double* __restrict__ partial )
{
#pragma clang fp contract( off )
    __shared__ double red[256];
    const PDArgs& A = *specs;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = A.n;
    double mx = 0.0;
    if ( p < n )
    {
        const double half_dt_m = A.half_dt / A.rho[p];
        const double* u = A.U[parity];
        double* un = A.U[parity ^ 1];
        double dd[3];
        for ( int d = 0; d < 3; d++ )
        {
            double vv = A.v[d * n + p];
            vv = vv + half_dt_m * A.f[d * n + p];
            A.v[d * n + p] = vv;
            const double nu = u[d * n + p] + A.dt * vv;
            un[d * n + p] = nu;
            dd[d] = nu - A.u_neigh[d * n + p];
        }
        mx = sqrt( dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] );
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for ( int w = 128; w > 0; w >>= 1 )
    {
        if ( (int)threadIdx.x < w )
            red[threadIdx.x] = fmax( red[threadIdx.x], red[threadIdx.x + w] );
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        partial[blockIdx.x] = red[0];
}

// Maximum of the per-block maxima, written to out (one block).
__global__ __launch_bounds__( 256 ) void kMaxReduce( const double* __restrict__ partial, int64_t nb,
                                                     double* __restrict__ out )
{
    __shared__ double red[256];
    double mx = 0.0;
    for ( int64_t b = threadIdx.x; b < nb; b += 256 )
        mx = fmax( mx, partial[b] );
    red[threadIdx.x] = mx;
    __syncthreads();
    for ( int w = 128; w > 0; w >>= 1 )
    {
        if ( (int)threadIdx.x < w )
            red[threadIdx.x] = fmax( red[threadIdx.x], red[threadIdx.x + w] );
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        *out = red[0];
}

using ContactFn = void ( * )( const PDArgs*, int32_t, int32_t );
ContactFn pickContact( const PDArgs& a )
{
    const bool jkr = a.model == PDK_MODEL_HERTZ_JKR;
    if ( a.polydisperse )
        return jkr ? kContact<true, true> : kContact<true, false>;
    return jkr ? kContact<false, true> : kContact<false, false>;
}

} // namespace

const char* contactKernelName( const PDArgs& a )
{
    const bool jkr = a.model == PDK_MODEL_HERTZ_JKR;
    if ( a.polydisperse )
        return jkr ? "kContact<true, true>" : "kContact<true, false>";
    return jkr ? "kContact<false, true>" : "kContact<false, false>";
}

void launchContactForce( const PDArgs* args_dev, const PDArgs& a, int32_t epilogue_mode, int32_t parity,
                         hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    const unsigned g = static_cast<unsigned>( ( a.nslices + kSlicesPerBlock - 1 ) / kSlicesPerBlock );
    hipLaunchKernelGGL( pickContact( a ), dim3( g ), dim3( kBlock ), 0, st, args_dev, epilogue_mode, parity );
    PDK_HIP( hipGetLastError() );
}

void launchContactInitialHalfStep( const PDArgs* args_dev, const PDArgs& a, int32_t parity, double* partial_dev,
                                   int64_t nblocks, hipStream_t st )
{
    hipLaunchKernelGGL( kContactInitial, dim3( static_cast<unsigned>( nblocks ) ), dim3( 256 ), 0, st, args_dev,
                        parity, partial_dev );
    PDK_HIP( hipGetLastError() );
}

void contactBuild( pdk_spec_s* s, bool initial )
{
    ContactState cs = contactState( s );
    contactBuildImpl( s, initial, cs.max_disp, cs.since_build );
}

// Neighbor::update then the contact pass, storing f (init / computeForce).
void contactForceStore( pdk_spec_s* s )
{
    ContactState cs = contactState( s );
    SpecView v = view( s );
    if ( *cs.max_disp > cs.extend )
        specTimed( s, 3, [&] { contactBuildImpl( s, false, cs.max_disp, cs.since_build ); } );
    v = view( s );
    specTimed( s, 0, [&] { launchContactForce( v.dargs, *v.h, kEpiStoreOnly, *v.cur, v.st ); } );
}

void contactStep( pdk_spec_s* s, int32_t nsteps )
{
    ContactState cs = contactState( s );
    for ( int32_t k = 0; k < nsteps; k++ )
    {
        SpecView v = view( s );
        const int64_t nb = ( v.n + 255 ) / 256;
        if ( cs.partial->bytes() < static_cast<size_t>( nb + 1 ) * sizeof( double ) )
            cs.partial->alloc( ( nb + 1 ) * sizeof( double ), cs.acct );
        double* part = cs.partial->as<double>();
        specTimed( s, 4, [&] {
            launchContactInitialHalfStep( v.dargs, *v.h, *v.cur, part, nb, v.st );
            hipLaunchKernelGGL( kMaxReduce, dim3( 1 ), dim3( 256 ), 0, v.st, part, nb, part + nb );
            PDK_HIP( hipGetLastError() );
        } );
        *v.cur ^= 1;
        // the host decides on a rebuild (CabanaPD_Neighbor.hpp:116-119)
        double mx = 0.0;
        PDK_HIP( hipMemcpyAsync( &mx, part + nb, sizeof( double ), hipMemcpyDeviceToHost, v.st ) );
        PDK_HIP( hipStreamSynchronize( v.st ) );
        *cs.max_disp = mx;
        *cs.since_build += 1;
        const bool periodic = cs.rebuild_every > 0 && *cs.since_build >= cs.rebuild_every;
        if ( mx > cs.extend || periodic )
            specTimed( s, 3, [&] { contactBuildImpl( s, false, cs.max_disp, cs.since_build ); } );
        v = view( s );
        specTimed( s, 0, [&] { launchContactForce( v.dargs, *v.h, kEpiFinal, *v.cur, v.st ); } );
    }
}

} // namespace pdk
