// pdk_build.hip -- device construction of the bond family.
//
// Replaces the reference's Cabana::VerletList build (called at
// CabanaPD_Neighbor.hpp:56-64 and :123-125) and Prenotch<N>::create
// (CabanaPD_Prenotch.hpp:215-263).  The set semantics are the reference's:
// every j != i with |p_j - p_i|^2 <= rc^2 (full list).  The layout is ours:
// particles in Morton order, rows in 64-row slices (one wavefront each),
// neighbors in ascending caller id, one intact bit per bond.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pdk_internal.hpp"

namespace pdk
{
namespace
{
constexpr int kThreads = 256;

inline unsigned gridFor( int64_t n, int t = kThreads )
{
    return static_cast<unsigned>( ( n + t - 1 ) / t );
}

__global__ void kMortonKeys( const double* __restrict__ xa, int64_t n, double lo0, double lo1,
                             double lo2, double inv_h, uint64_t* __restrict__ keys,
                             int64_t* __restrict__ vals )
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( c >= n )
        return;
    const double lo[3] = { lo0, lo1, lo2 };
    uint64_t q[3];
    for ( int d = 0; d < 3; d++ )
    {
        double t = ( xa[3 * c + d] - lo[d] ) * inv_h + 0.5;
        t = fmin( fmax( t, 0.0 ), 2097151.0 );
        q[d] = static_cast<uint64_t>( t );
    }
    keys[c] = morton3( q[0], q[1], q[2] );
    vals[c] = c;
}

__global__ void kInvertPerm( const int64_t* __restrict__ perm, int64_t n, int64_t* __restrict__ inv )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p < n )
        inv[perm[p]] = p;
}

__global__ void kGatherAos3( const double* __restrict__ src, const int64_t* __restrict__ perm,
                             double* __restrict__ dst, int64_t n )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p >= n )
        return;
    const int64_t c = perm[p];
    dst[p] = src[3 * c + 0];
    dst[n + p] = src[3 * c + 1];
    dst[2 * n + p] = src[3 * c + 2];
}

__global__ void kScatterAos3( const double* __restrict__ src, const int64_t* __restrict__ perm,
                              double* __restrict__ dst, int64_t n )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p >= n )
        return;
    const int64_t c = perm[p];
    dst[3 * c + 0] = src[p];
    dst[3 * c + 1] = src[n + p];
    dst[3 * c + 2] = src[2 * n + p];
}

__global__ void kGatherScalar( const double* __restrict__ src, const int64_t* __restrict__ perm,
                               double* __restrict__ dst, int64_t n )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p < n )
        dst[p] = src[perm[p]];
}

__global__ void kScatterScalar( const double* __restrict__ src, const int64_t* __restrict__ perm,
                                double* __restrict__ dst, int64_t n )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p < n )
        dst[perm[p]] = src[p];
}

__global__ void kGatherFlags( const int32_t* __restrict__ src, const int64_t* __restrict__ perm,
                              uint8_t* __restrict__ dst, int64_t n )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p < n )
        dst[p] = src[perm[p]] != 0 ? 1 : 0;
}

__device__ __forceinline__ void posOf( const NeighborQuery& q, int64_t i, double& a, double& b,
                                       double& c )
{
#pragma clang fp contract( off )
    a = q.x[i];
    b = q.x[q.n + i];
    c = q.x[2 * q.n + i];
    if ( q.u )
    {
        // y = x + u (CabanaPD_Particles.hpp:543-551)
        a = a + q.u[i];
        b = b + q.u[q.n + i];
        c = c + q.u[2 * q.n + i];
    }
}

// Per-block min/max of the query positions.
__global__ void kBBox( NeighborQuery q, double* __restrict__ partial )
{
    __shared__ double red[6][kThreads];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double p[3] = { 1e300, 1e300, 1e300 };
    double h[3] = { -1e300, -1e300, -1e300 };
    if ( i < q.n )
    {
        posOf( q, i, p[0], p[1], p[2] );
        h[0] = p[0];
        h[1] = p[1];
        h[2] = p[2];
    }
    for ( int d = 0; d < 3; d++ )
    {
        red[d][threadIdx.x] = p[d];
        red[3 + d][threadIdx.x] = h[d];
    }
    __syncthreads();
    for ( int w = kThreads / 2; w > 0; w >>= 1 )
    {
        if ( (int)threadIdx.x < w )
            for ( int d = 0; d < 3; d++ )
            {
                red[d][threadIdx.x] = fmin( red[d][threadIdx.x], red[d][threadIdx.x + w] );
                red[3 + d][threadIdx.x] = fmax( red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + w] );
            }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        for ( int d = 0; d < 6; d++ )
            partial[6 * blockIdx.x + d] = red[d][0];
}

struct CellGrid
{
    double lo[3];
    double ext[3];
    int32_t nc[3];
};

__device__ __forceinline__ int32_t cellCoord( const CellGrid& g, double p, int d )
{
    int32_t c = 0;
    if ( g.ext[d] > 0 )
    {
        double t = ( p - g.lo[d] ) / g.ext[d] * g.nc[d];
        c = static_cast<int32_t>( fmin( fmax( t, 0.0 ), (double)( g.nc[d] - 1 ) ) );
    }
    return c;
}

__global__ void kCellIds( NeighborQuery q, CellGrid g, int32_t* __restrict__ cell,
                          int32_t* __restrict__ count )
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( i >= q.n )
        return;
    double a, b, c;
    posOf( q, i, a, b, c );
    const int32_t id = ( cellCoord( g, c, 2 ) * g.nc[1] + cellCoord( g, b, 1 ) ) * g.nc[0] + cellCoord( g, a, 0 );
    cell[i] = id;
    atomicAdd( &count[id], 1 );
}

__global__ void kCellFill( const int32_t* __restrict__ cell, int64_t n, int32_t* __restrict__ fill,
                           int32_t* __restrict__ list )
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    const int32_t pos = atomicAdd( &fill[cell[i]], 1 );
    list[pos] = static_cast<int32_t>( i );
}

// Visit all j != i within the cutoff; fn(j) is called in cell order.
template <class Fn>
__device__ __forceinline__ void visitNeighbors( const NeighborQuery& q, const CellGrid& g,
                                                const int32_t* __restrict__ cell_start,
                                                const int32_t* __restrict__ list, int64_t i, Fn&& fn )
{
#pragma clang fp contract( off )
    double pi0, pi1, pi2;
    posOf( q, i, pi0, pi1, pi2 );
    const int32_t cx = cellCoord( g, pi0, 0 ), cy = cellCoord( g, pi1, 1 ), cz = cellCoord( g, pi2, 2 );
    const double rc2 = q.rc * q.rc;
    const double ri = q.radius ? q.radius[i] : 0.0;
    for ( int32_t z = max( 0, cz - 1 ); z <= min( g.nc[2] - 1, cz + 1 ); z++ )
        for ( int32_t y = max( 0, cy - 1 ); y <= min( g.nc[1] - 1, cy + 1 ); y++ )
            for ( int32_t x = max( 0, cx - 1 ); x <= min( g.nc[0] - 1, cx + 1 ); x++ )
            {
                const int32_t c = ( z * g.nc[1] + y ) * g.nc[0] + x;
                for ( int32_t k = cell_start[c]; k < cell_start[c + 1]; k++ )
                {
                    const int64_t j = list[k];
                    if ( j == i )
                        continue;
                    double pj0, pj1, pj2;
                    posOf( q, j, pj0, pj1, pj2 );
                    const double dx = pj0 - pi0;
                    const double dy = pj1 - pi1;
                    const double dz = pj2 - pi2;
                    const double dist2 = dx * dx + dy * dy + dz * dz;
                    double cut2 = rc2;
                    if ( q.radius )
                    {
                        const double cut = ri + q.radius[j] + q.extend;
                        cut2 = cut * cut;
                    }
                    if ( dist2 <= cut2 )
                        fn( j );
                }
            }
}

// One wavefront per slice: per-row counts and the slice's slot-row count.
__global__ __launch_bounds__( kThreads ) void kCountNeighbors( NeighborQuery q, CellGrid g,
                                                               const int32_t* __restrict__ cell_start,
                                                               const int32_t* __restrict__ list,
                                                               int32_t* __restrict__ cnt,
                                                               int64_t* __restrict__ slice_len )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int32_t c = 0;
    if ( p < q.n )
        visitNeighbors( q, g, cell_start, list, p, [&]( int64_t ) { c++; } );
    if ( p < q.n )
        cnt[p] = c;
    // wavefront max
    int32_t m = c;
    for ( int off = 32; off > 0; off >>= 1 )
        m = max( m, __shfl_xor( m, off ) );
    const int64_t slice = p / kSlice;
    const int64_t nslices = ( q.n + kSlice - 1 ) / kSlice;
    if ( ( threadIdx.x & ( kSlice - 1 ) ) == 0 && slice < nslices )
        slice_len[slice] = m;
}

__global__ __launch_bounds__( kThreads ) void kFillNeighbors( NeighborQuery q, CellGrid g,
                                                              const int32_t* __restrict__ cell_start,
                                                              const int32_t* __restrict__ list,
                                                              const int64_t* __restrict__ perm,
                                                              const int64_t* __restrict__ inv,
                                                              const int64_t* __restrict__ slice_ptr,
                                                              const int32_t* __restrict__ cnt,
                                                              int32_t* __restrict__ col,
                                                              uint64_t* __restrict__ intact )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t slice = p / kSlice;
    const int64_t nslices = ( q.n + kSlice - 1 ) / kSlice;
    if ( slice >= nslices )
        return; // whole wavefront
    const int64_t base = slice_ptr[slice];
    const int32_t len = static_cast<int32_t>( slice_ptr[slice + 1] - base );
    int32_t* row = col + base * kSlice + lane;
    const bool valid = p < q.n;
    int32_t c = 0;
    if ( valid )
    {
        // Collect caller ids, insertion-sort them, map back to sorted ids.
        visitNeighbors( q, g, cell_start, list, p, [&]( int64_t j ) {
            if ( c >= len )
                return; // cannot happen: len is the slice's maximum count
            const int32_t key = static_cast<int32_t>( perm[j] );
            int32_t k = c;
            while ( k > 0 && row[( k - 1 ) * kSlice] > key )
            {
                row[k * kSlice] = row[( k - 1 ) * kSlice];
                k--;
            }
            row[k * kSlice] = key;
            c++;
        } );
        for ( int32_t k = 0; k < c; k++ )
            row[k * kSlice] = static_cast<int32_t>( inv[row[k * kSlice]] );
    }
    const int32_t self = static_cast<int32_t>( valid ? p : slice * kSlice );
    for ( int32_t k = c; k < len; k++ )
        row[k * kSlice] = self;
    const int32_t mycnt = valid ? cnt[p] : 0;
    for ( int32_t k = 0; k < len; k++ )
    {
        const uint64_t bits = __ballot( k < mycnt );
        if ( lane == 0 )
            intact[base + k] = bits;
    }
}

// ---- Prenotch (CabanaPD_Prenotch.hpp:24-172), FP64 op for op --------------
struct D3
{
    double a[3];
};
__device__ __forceinline__ double pdot( D3 a, D3 b )
{
#pragma clang fp contract( off )
    double v = 0;
    for ( int d = 0; d < 3; d++ )
        v += ( a.a[d] ) * ( b.a[d] );
    return v;
}
__device__ __forceinline__ D3 pcross( D3 a, D3 b )
{
#pragma clang fp contract( off )
    D3 c;
    c.a[0] = a.a[1] * b.a[2] - a.a[2] * b.a[1];
    c.a[1] = -( a.a[0] * b.a[2] - a.a[2] * b.a[0] );
    c.a[2] = a.a[0] * b.a[1] - a.a[1] * b.a[0];
    return c;
}
__device__ __forceinline__ D3 pscale( D3 a, double s )
{
#pragma clang fp contract( off )
    for ( int d = 0; d < 3; d++ )
        a.a[d] *= s;
    return a;
}
__device__ __forceinline__ D3 pdiff( D3 a, D3 b )
{
#pragma clang fp contract( off )
    D3 c;
    for ( int d = 0; d < 3; d++ )
        c.a[d] = a.a[d] - b.a[d];
    return c;
}
__device__ __forceinline__ D3 psum( D3 a, D3 b )
{
#pragma clang fp contract( off )
    D3 c;
    for ( int d = 0; d < 3; d++ )
        c.a[d] = a.a[d] + b.a[d];
    return c;
}

__device__ int bondPrenotchIntersection( D3 v1, D3 v2, D3 p0, D3 x_i, D3 x_j )
{
#pragma clang fp contract( off )
    const double tol = 1e-10;
    D3 cross_v1_v2 = pcross( v1, v2 );
    double norm_cross_v1_v2 = sqrt( pdot( cross_v1_v2, cross_v1_v2 ) );
    D3 nrm = pscale( cross_v1_v2, 1.0 / norm_cross_v1_v2 );
    D3 l0 = x_i;
    D3 l = pdiff( x_j, x_i );
    int case_flag = 3;
    if ( fabs( pdot( l, nrm ) ) < tol )
        case_flag = ( fabs( pdot( pdiff( p0, l0 ), nrm ) ) < tol ) ? 1 : 2;
    int keep_bond = 1;
    if ( case_flag == 1 )
    {
        double norm2 = norm_cross_v1_v2 * norm_cross_v1_v2;
        double li1 = pdot( pcross( pdiff( x_i, p0 ), v2 ), cross_v1_v2 ) / norm2;
        double li2 = -pdot( pcross( pdiff( x_i, p0 ), v1 ), cross_v1_v2 ) / norm2;
        double lj1 = pdot( pcross( pdiff( x_j, p0 ), v2 ), cross_v1_v2 ) / norm2;
        double lj2 = -pdot( pcross( pdiff( x_j, p0 ), v1 ), cross_v1_v2 ) / norm2;
        if ( !( fmin( li1, lj1 ) > 1 + tol || fmax( li1, lj1 ) < -tol || fmin( li2, lj2 ) > 1 + tol ||
                fmax( li2, lj2 ) < -tol ) )
            keep_bond = 0;
    }
    else if ( case_flag == 3 )
    {
        double dd = pdot( pdiff( p0, l0 ), nrm ) / pdot( l, nrm );
        if ( -tol < dd && dd < 1 + tol )
        {
            D3 p = psum( l0, pscale( l, dd ) );
            double norm2 = norm_cross_v1_v2 * norm_cross_v1_v2;
            double l1 = pdot( pcross( pdiff( p, p0 ), v2 ), cross_v1_v2 ) / norm2;
            double l2 = -pdot( pcross( pdiff( p, p0 ), v1 ), cross_v1_v2 ) / norm2;
            if ( -tol < fmin( l1, l2 ) && fmax( l1, l2 ) < 1 + tol )
                keep_bond = 0;
        }
    }
    return keep_bond;
}

__global__ __launch_bounds__( kThreads ) void kPrenotch( PDArgs a, int32_t nnotch,
                                                         const double* __restrict__ notch )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t slice = p / kSlice;
    if ( slice >= a.nslices )
        return;
    const bool valid = p < a.n;
    const int64_t pc = valid ? p : slice * kSlice;
    const int64_t base = a.slice_ptr[slice];
    const int32_t len = static_cast<int32_t>( a.slice_ptr[slice + 1] - base );
    const int32_t mycnt = valid ? a.cnt[p] : 0;
    const int64_t n = a.n;
    D3 xi{ { a.x[pc], a.x[n + pc], a.x[2 * n + pc] } };
    for ( int32_t k = 0; k < len; k++ )
    {
        bool cut = false;
        if ( k < mycnt )
        {
            const int64_t j = a.col[( base + k ) * kSlice + lane];
            D3 xj{ { a.x[j], a.x[n + j], a.x[2 * n + j] } };
            for ( int32_t t = 0; t < nnotch; t++ )
            {
                const double* nt = notch + 9 * t;
                D3 p0{ { nt[0], nt[1], nt[2] } }, v1{ { nt[3], nt[4], nt[5] } }, v2{ { nt[6], nt[7], nt[8] } };
                if ( !bondPrenotchIntersection( v1, v2, p0, xi, xj ) )
                    cut = true;
            }
        }
        const uint64_t m = __ballot( cut );
        if ( m && lane == 0 )
            a.intact[base + k] &= ~m;
    }
}

template <class T>
void ensure( DevBuf& b, size_t count, MemAccount* acct )
{
    if ( b.bytes() < count * sizeof( T ) )
        b.alloc( std::max<size_t>( count * sizeof( T ), 256 ), acct );
}

} // namespace

void launchMortonOrder( const double* x_aos_dev, int64_t n, const double lo[3], double inv_h,
                        int64_t* perm_dev, int64_t* inv_dev, BuildScratch& s, MemAccount* acct,
                        hipStream_t st )
{
    ensure<uint64_t>( s.keys, n, acct );
    ensure<uint64_t>( s.keys2, n, acct );
    ensure<int64_t>( s.vals, n, acct );
    kMortonKeys<<<gridFor( n ), kThreads, 0, st>>>( x_aos_dev, n, lo[0], lo[1], lo[2], inv_h,
                                                    s.keys.as<uint64_t>(), s.vals.as<int64_t>() );
    PDK_HIP( hipGetLastError() );
    size_t tmp = 0;
    PDK_HIP( hipcub::DeviceRadixSort::SortPairs( nullptr, tmp, s.keys.as<uint64_t>(), s.keys2.as<uint64_t>(),
                                                 s.vals.as<int64_t>(), perm_dev, static_cast<int>( n ), 0, 63,
                                                 st ) );
    ensure<char>( s.tmp, tmp, acct );
    PDK_HIP( hipcub::DeviceRadixSort::SortPairs( s.tmp.as<void>(), tmp, s.keys.as<uint64_t>(),
                                                 s.keys2.as<uint64_t>(), s.vals.as<int64_t>(), perm_dev,
                                                 static_cast<int>( n ), 0, 63, st ) );
    kInvertPerm<<<gridFor( n ), kThreads, 0, st>>>( perm_dev, n, inv_dev );
    PDK_HIP( hipGetLastError() );
}

void launchGatherAos3( const double* src_aos, const int64_t* perm, double* dst_soa, int64_t n, hipStream_t st )
{
    if ( n == 0 )
        return;
    kGatherAos3<<<gridFor( n ), kThreads, 0, st>>>( src_aos, perm, dst_soa, n );
    PDK_HIP( hipGetLastError() );
}
void launchScatterAos3( const double* src_soa, const int64_t* perm, double* dst_aos, int64_t n, hipStream_t st )
{
    if ( n == 0 )
        return;
    kScatterAos3<<<gridFor( n ), kThreads, 0, st>>>( src_soa, perm, dst_aos, n );
    PDK_HIP( hipGetLastError() );
}
void launchGatherScalar( const double* src, const int64_t* perm, double* dst, int64_t n, hipStream_t st )
{
    if ( n == 0 )
        return;
    kGatherScalar<<<gridFor( n ), kThreads, 0, st>>>( src, perm, dst, n );
    PDK_HIP( hipGetLastError() );
}
void launchScatterScalar( const double* src, const int64_t* perm, double* dst, int64_t n, hipStream_t st )
{
    if ( n == 0 )
        return;
    kScatterScalar<<<gridFor( n ), kThreads, 0, st>>>( src, perm, dst, n );
    PDK_HIP( hipGetLastError() );
}
void launchGatherFlags( const int32_t* src, const int64_t* perm, uint8_t* dst, int64_t n, hipStream_t st )
{
    if ( n == 0 )
        return;
    kGatherFlags<<<gridFor( n ), kThreads, 0, st>>>( src, perm, dst, n );
    PDK_HIP( hipGetLastError() );
}

int64_t buildSlicedCsr( const NeighborQuery& q, const int64_t* perm, const int64_t* inv,
                        int64_t* slice_ptr, int32_t* cnt, DevBuf& col, DevBuf& intact,
                        int64_t& col_capacity_rows, BuildScratch& s, MemAccount* acct,
                        hipStream_t st, int32_t* max_cnt_out )
{
    const int64_t n = q.n;
    const int64_t nslices = ( n + kSlice - 1 ) / kSlice;
    // Bounding box.
    const int64_t nb = gridFor( n );
    ensure<double>( s.bbox, 6 * nb, acct );
    kBBox<<<static_cast<unsigned>( nb ), kThreads, 0, st>>>( q, s.bbox.as<double>() );
    PDK_HIP( hipGetLastError() );
    std::vector<double> part( 6 * nb );
    PDK_HIP( hipMemcpyAsync( part.data(), s.bbox.as<double>(), 6 * nb * sizeof( double ), hipMemcpyDeviceToHost, st ) );
    PDK_HIP( hipStreamSynchronize( st ) );
    double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
    for ( int64_t b = 0; b < nb; b++ )
        for ( int d = 0; d < 3; d++ )
        {
            lo[d] = std::min( lo[d], part[6 * b + d] );
            hi[d] = std::max( hi[d], part[6 * b + 3 + d] );
        }
    // Cell grid with cells no smaller than the largest pair cutoff.
    double cell = q.rc;
    if ( q.radius )
    {
        std::vector<double> rh( n );
        PDK_HIP( hipMemcpyAsync( rh.data(), q.radius, n * sizeof( double ), hipMemcpyDeviceToHost, st ) );
        PDK_HIP( hipStreamSynchronize( st ) );
        double rmax = 0;
        for ( double r : rh )
            rmax = std::max( rmax, r );
        cell = 2.0 * rmax + q.extend;
    }
    CellGrid g;
    int64_t ncell = 1;
    for ( int d = 0; d < 3; d++ )
    {
        g.lo[d] = lo[d];
        g.ext[d] = hi[d] - lo[d];
        int64_t c = std::max<int64_t>( 1, static_cast<int64_t>( g.ext[d] / cell ) );
        c = std::min<int64_t>( c, 1024 );
        g.nc[d] = static_cast<int32_t>( c );
        ncell *= c;
    }
    ensure<int32_t>( s.cell, n, acct );
    ensure<int32_t>( s.cell_start, ncell + 1, acct );
    ensure<int32_t>( s.cell_fill, ncell + 1, acct );
    ensure<int32_t>( s.cell_list, n, acct );
    PDK_HIP( hipMemsetAsync( s.cell_fill.as<int32_t>(), 0, ( ncell + 1 ) * sizeof( int32_t ), st ) );
    kCellIds<<<gridFor( n ), kThreads, 0, st>>>( q, g, s.cell.as<int32_t>(), s.cell_fill.as<int32_t>() );
    PDK_HIP( hipGetLastError() );
    size_t tmp = 0;
    PDK_HIP( hipcub::DeviceScan::ExclusiveSum( nullptr, tmp, s.cell_fill.as<int32_t>(), s.cell_start.as<int32_t>(),
                                               static_cast<int>( ncell + 1 ), st ) );
    ensure<char>( s.tmp, tmp, acct );
    PDK_HIP( hipcub::DeviceScan::ExclusiveSum( s.tmp.as<void>(), tmp, s.cell_fill.as<int32_t>(),
                                               s.cell_start.as<int32_t>(), static_cast<int>( ncell + 1 ), st ) );
    PDK_HIP( hipMemcpyAsync( s.cell_fill.as<int32_t>(), s.cell_start.as<int32_t>(), ( ncell + 1 ) * sizeof( int32_t ),
                             hipMemcpyDeviceToDevice, st ) );
    kCellFill<<<gridFor( n ), kThreads, 0, st>>>( s.cell.as<int32_t>(), n, s.cell_fill.as<int32_t>(),
                                                  s.cell_list.as<int32_t>() );
    PDK_HIP( hipGetLastError() );
    // Counts and slot rows per slice.
    ensure<int64_t>( s.slice_len, nslices + 1, acct );
    PDK_HIP( hipMemsetAsync( s.slice_len.as<int64_t>(), 0, ( nslices + 1 ) * sizeof( int64_t ), st ) );
    const unsigned gslices = static_cast<unsigned>( ( nslices * kSlice + kThreads - 1 ) / kThreads );
    kCountNeighbors<<<gslices, kThreads, 0, st>>>( q, g, s.cell_start.as<int32_t>(), s.cell_list.as<int32_t>(), cnt,
                                                    s.slice_len.as<int64_t>() );
    PDK_HIP( hipGetLastError() );
    tmp = 0;
    PDK_HIP( hipcub::DeviceScan::ExclusiveSum( nullptr, tmp, s.slice_len.as<int64_t>(), slice_ptr,
                                               static_cast<int>( nslices + 1 ), st ) );
    ensure<char>( s.tmp, tmp, acct );
    PDK_HIP( hipcub::DeviceScan::ExclusiveSum( s.tmp.as<void>(), tmp, s.slice_len.as<int64_t>(), slice_ptr,
                                               static_cast<int>( nslices + 1 ), st ) );
    int64_t total_rows = 0;
    PDK_HIP( hipMemcpyAsync( &total_rows, slice_ptr + nslices, sizeof( int64_t ), hipMemcpyDeviceToHost, st ) );
    PDK_HIP( hipStreamSynchronize( st ) );
    if ( total_rows > col_capacity_rows || !col )
    {
        const int64_t cap = std::max<int64_t>( total_rows + total_rows / 8, 1 );
        col.alloc( static_cast<size_t>( cap ) * kSlice * sizeof( int32_t ), acct );
        intact.alloc( static_cast<size_t>( cap ) * sizeof( uint64_t ), acct );
        col_capacity_rows = cap;
    }
    kFillNeighbors<<<gslices, kThreads, 0, st>>>( q, g, s.cell_start.as<int32_t>(), s.cell_list.as<int32_t>(), perm,
                                                   inv, slice_ptr, cnt, col.as<int32_t>(), intact.as<uint64_t>() );
    PDK_HIP( hipGetLastError() );
    if ( max_cnt_out )
    {
        std::vector<int64_t> sl( nslices );
        PDK_HIP( hipMemcpyAsync( sl.data(), s.slice_len.as<int64_t>(), nslices * sizeof( int64_t ),
                                 hipMemcpyDeviceToHost, st ) );
        PDK_HIP( hipStreamSynchronize( st ) );
        int64_t m = 0;
        for ( int64_t v : sl )
            m = std::max( m, v );
        *max_cnt_out = static_cast<int32_t>( m );
    }
    return total_rows;
}

void launchPrenotch( const PDArgs& a, int32_t nnotch, const double* notch_dev, hipStream_t st )
{
    const unsigned g = static_cast<unsigned>( ( a.nslices * kSlice + kThreads - 1 ) / kThreads );
    kPrenotch<<<g, kThreads, 0, st>>>( a, nnotch, notch_dev );
    PDK_HIP( hipGetLastError() );
}

} // namespace pdk
