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

// Correctly rounded powers of the contact models (the reference evaluates
// Kokkos::pow, within an ulp of the exact value).  Built from IEEE operations
// only -- + - * /, sqrt, fma, frexp/ldexp -- in the same order as the
// oracle's, so both return the same bits.  Double-double arithmetic (exact
// products through fma, after Dekker) carries each power to ~2^-100 relative
// before the final rounding, which is then the correct rounding unless the
// exact value lies within ~2^-100 of a rounding boundary (Muller et al.,
// Handbook of Floating-Point Arithmetic, on correct rounding).
__device__ __forceinline__ double crSq( double x )
{
    return x * x;
}
__device__ double crCube( double x )
{
#pragma clang fp contract( off )
    if ( !( fabs( x ) < 1e100 ) || x == 0.0 )
        return x * x * x; // overflow, infinities, NaN and signed zeros as pow
    const double ph = x * x, pl = fma( x, x, -ph );
    const double qh = ph * x, ql = fma( ph, x, -qh );
    return qh + ( ql + pl * x );
}
__device__ __forceinline__ double crHalf( double x )
{
    return x == 0.0 ? 0.0 : sqrt( x ); // +0 at -0, as pow
}
// x^1.5: sqrt with its exact remainder x - sq^2, times x in double-double.
__device__ double crPow32( double x )
{
#pragma clang fp contract( off )
    if ( !( x > 0.0 ) )
        return x == 0.0 ? 0.0 : sqrt( -1.0 );
    if ( isinf( x ) )
        return x;
    const double sq = sqrt( x );
    const double rem = fma( -sq, sq, x );
    const double sl = rem / ( 2.0 * sq );
    const double ph = x * sq, pl = fma( x, sq, -ph );
    return ph + ( pl + x * sl );
}
// ln x for x > 0 to ~1e-16 relative: e ln 2 + 2 atanh(z), z = (m-1)/(m+1),
// m in [1/sqrt 2, sqrt 2), the atanh series to z^23.
__device__ double crLog( double x )
{
#pragma clang fp contract( off )
    int e = 0;
    double m = frexp( x, &e );
    if ( m < 0.70710678118654752440 )
    {
        m = 2.0 * m;
        e = e - 1;
    }
    const double z = ( m - 1.0 ) / ( m + 1.0 );
    const double z2 = z * z;
    double sum = 1.0 / 23.0;
    for ( int k = 21; k >= 1; k -= 2 )
        sum = sum * z2 + 1.0 / k;
    return static_cast<double>( e ) * 0.69314718055994530942 + 2.0 * z * sum;
}
// x^y with y the double nearest 1/3, y = (1 - 2^-54) / 3: cbrt(x) as a
// double-double (Newton from a cubic start, the last step with the exact
// double-double cube), times x^(-2^-54/3) = 1 - t + t^2/2, t = ln(x)/(3 2^54).
__device__ double crThird( double x )
{
#pragma clang fp contract( off )
    if ( x != x || x < 0.0 )
        return sqrt( -1.0 );
    if ( x == 0.0 )
        return 0.0;
    if ( isinf( x ) )
        return x;
    int e = 0;
    const double m0 = frexp( x, &e ); // x = m0 2^e, m0 in [0.5, 1)
    int r = e % 3;
    if ( r < 0 )
        r += 3;
    const double m = ldexp( m0, r ); // in [0.5, 4), e - r divisible by 3
    const int q = ( e - r ) / 3;
    // Newton for c^3 = m from a cubic start (cbrt on [0.5, 1), relative
    // error below 1.4e-4, times 2^(r/3)): three steps reach ~1 ulp
    double c = 0.44019728644131806 + m0 * ( 0.9266652504745495 + m0 * ( -0.5110337078357359 + m0 * 0.1442771968559388 ) );
    c = c * ( r == 0 ? 1.0 : ( r == 1 ? 1.2599210498948732 : 1.5874010519681994 ) );
    for ( int it = 0; it < 3; it++ )
        c = c - ( c * c * c - m ) / ( 3.0 * c * c );
    const double ph = c * c, pl = fma( c, c, -ph );
    const double qh = ph * c, ql = fma( ph, c, -qh );
    const double res = ( ( m - qh ) - ql ) - pl * c;
    const double cl = res / ( 3.0 * ph );
    const double t = crLog( x ) * ( 1.0 / 54043195528445952.0 );
    const double k = -t + 0.5 * ( t * t );
    const double chi = ldexp( c, q ), clo = ldexp( cl, q );
    return chi + ( clo + chi * k );
}

// cbrt(x) to 1.4e-4 relative for finite x > 0: the cubic start of crThird
// (the exponent (1 - 2^-54) / 3 of the exact power changes the value by less
// than 1.3e-14 relative, which the bound absorbs).
__device__ __forceinline__ double cbrtEstimate( double x )
{
#pragma clang fp contract( off )
    int e = 0;
    const double m0 = frexp( x, &e );
    int r = e % 3;
    if ( r < 0 )
        r += 3;
    const int q = ( e - r ) / 3;
    double c = 0.44019728644131806 + m0 * ( 0.9266652504745495 + m0 * ( -0.5110337078357359 + m0 * 0.1442771968559388 ) );
    c = c * ( r == 0 ? 1.0 : ( r == 1 ? 1.2599210498948732 : 1.5874010519681994 ) );
    return ldexp( c, q );
}

__device__ double jkrPatchRadius( const PDArgs& A, double Rs, double delta_n )
{
#pragma clang fp contract( off )
    // force_models/CabanaPD_HertzianJKR.hpp:141-176
    const double nu = A.nu, E = A.E, Gamma = A.Gamma;
    double c0 = crSq( Rs * delta_n );
    double c1 = -4.0 * ( 1.0 - crSq( nu ) ) * kPi * Gamma * crSq( Rs ) / E;
    double c2 = -2.0 * Rs * delta_n;
    double P = -crSq( c2 ) / 12.0 - c0;
    double Q = -crCube( c2 ) / 108.0 + c0 * c2 / 3.0 - crSq( c1 ) / 8.0;
    double U = crThird( -Q / 2.0 + crHalf( fmax( 0.0, crSq( Q ) / 4.0 + crCube( P ) / 27.0 ) ) );
    double sv = 0.0;
    if ( P != 0.0 )
        sv = -5.0 / 6.0 * c2 + U - P / ( 3.0 * U );
    else
        sv = -5.0 / 6.0 * c2 - crThird( Q );
    double w = crHalf( fmax( 0.0, c2 + 2.0 * sv ) );
    double lambda = c1 / ( 2.0 * w + 1e-14 );
    return 1.0 / 2.0 * ( w + crHalf( fmax( 0.0, crSq( w ) - 4.0 * ( c2 + sv + lambda ) ) ) );
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
            coeff = fmin( 0.0, -pk.coeff_h_n * crPow32( fabs( delta_n ) ) );
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
            double a3 = crCube( a );
            coeff = -1.0 * ( pk.coeff_h_n * a3 - crHalf( 8.0 * A.Gamma * kPi * A.Es * a3 ) );
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
                const double xa = 9.0 / 2.0 * A.Gamma * kPi / A.Es * crSq( pk.Rs );
                // The tear distance only places delta_n = 2R - r on one side
                // of it (normalForce): an estimate within 3e-4 relative
                // decides unless delta_n lies within 1e-3 of it, and only
                // then is the correctly rounded value formed.  The decision,
                // and with it every force, is that of the exact value.
                // (pow(6, -1/3) of the host's libm, as the oracle evaluates it)
                const double a0e = cbrtEstimate( xa );
                const double dte = -1.0 / 2.0 * A.jkr_c6 * crSq( a0e ) / pk.Rs;
                const double dn = pk.two_r - r;
                if ( xa > 0.0 && xa < 1e300 && fabs( dte ) > 1e-290 && fabs( dte ) < 1e290 &&
                     fabs( dn - dte ) > 1e-3 * fabs( dte ) )
                    pk.delta_tear = dn >= dte ? -INFINITY : INFINITY;
                else
                {
                    const double a0 = crThird( xa );
                    pk.delta_tear = -1.0 / 2.0 * A.jkr_c6 * crSq( a0 ) / pk.Rs;
                }
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
    if ( epi != kEpiFinal )
        return;
    // final half step (CabanaPD_Integrate.hpp:128-134); only when the pair
    // force does not read velocities, since other waves of this launch read
    // v of their neighbors
    
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

// VelocityVerlet::finalHalfStep (CabanaPD_Integrate.hpp:128-134) as its own
// pass, after a contact pass whose pair forces read the velocities.
__global__ __launch_bounds__( 256 ) void kContactFinal( const PDArgs* __restrict__ specs )
{
#pragma clang fp contract( off )
    const PDArgs& A = *specs;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = A.n;
    if ( p >= n )
        return;
    const double half_dt_m = A.half_dt / A.rho[p];
    for ( int d = 0; d < 3; d++ )
        A.v[d * n + p] = A.v[d * n + p] + half_dt_m * A.f[d * n + p];
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

// Whether the pair force reads the velocities: the damping term (restitution
// below 1) and, for Hertz-JKR, the approach/recede switch and the friction
// (force_models/CabanaPD_Hertzian.hpp:74-80, CabanaPD_HertzianJKR.hpp:96-129).
// Undamped Hertz reads them only through a zero term.
bool contactReadsVelocity( const PDArgs& a )
{
    return a.model == PDK_MODEL_HERTZ_JKR || a.coeff_h_d != 0.0;
}

void launchContactForce( const PDArgs* args_dev, const PDArgs& a, int32_t epilogue_mode, int32_t parity,
                         hipStream_t st )
{
    if ( a.nslices == 0 )
        return;
    // a velocity-dependent pass must not update v while other waves read it
    if ( epilogue_mode == kEpiFinal && contactReadsVelocity( a ) )
        throw Error( "internal: velocity-dependent contact needs the split final half step" );
    const unsigned g = static_cast<unsigned>( ( a.nslices + kSlicesPerBlock - 1 ) / kSlicesPerBlock );
    hipLaunchKernelGGL( pickContact( a ), dim3( g ), dim3( kBlock ), 0, st, args_dev, epilogue_mode, parity );
    PDK_HIP( hipGetLastError() );
}

void launchContactFinal( const PDArgs* args_dev, const PDArgs& a, hipStream_t st )
{
    if ( a.n == 0 )
        return;
    hipLaunchKernelGGL( kContactFinal, dim3( static_cast<unsigned>( ( a.n + 255 ) / 256 ) ), dim3( 256 ), 0, st,
                        args_dev );
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
        // The host decides on a rebuild (CabanaPD_Neighbor.hpp:116-119) from
        // the largest displacement, copied to pinned memory.  Unless the
        // periodic rebuild is due, the force pass (which writes f only; the
        // final half step is its own pass) is launched with the current list
        // before the host waits for that copy, so the GPU is not left idle
        // meanwhile; if the displacement calls for a rebuild, the list is
        // rebuilt and the force pass runs again over the new list, which
        // overwrites f.  The results are those of deciding first.
        PDK_HIP( hipMemcpyAsync( cs.disp_pinned, part + nb, sizeof( double ), hipMemcpyDeviceToHost, v.st ) );
        PDK_HIP( hipEventRecord( cs.disp_ready, v.st ) );
        *cs.since_build += 1;
        const bool periodic = cs.rebuild_every > 0 && *cs.since_build >= cs.rebuild_every;
        auto force = [&] {
            v = view( s );
            specTimed( s, 0, [&] { launchContactForce( v.dargs, *v.h, kEpiApplyOnly, *v.cur, v.st ); } );
        };
        if ( !periodic )
            force();
        PDK_HIP( hipEventSynchronize( cs.disp_ready ) );
        const double mx = *cs.disp_pinned;
        *cs.max_disp = mx;
        if ( mx > cs.extend || periodic )
        {
            specTimed( s, 3, [&] { contactBuildImpl( s, false, cs.max_disp, cs.since_build ); } );
            force();
        }
        specTimed( s, 4, [&] { launchContactFinal( v.dargs, *v.h, v.st ); } );
    }
}

} // namespace pdk
