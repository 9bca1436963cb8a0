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

// The query position (x + u) and radius of every particle in cell-list
// order, so that a cell's candidates are one contiguous run of 32-byte
// records (one load each) instead of an id followed by dependent gathers.
__global__ void kCellPack( NeighborQuery q, const int32_t* __restrict__ list, double4* __restrict__ rec )
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( k >= q.n )
        return;
    const int64_t j = list[k];
    double a, b, c;
    posOf( q, j, a, b, c );
    rec[k] = make_double4( a, b, c, q.radius ? q.radius[j] : 0.0 );
}

// Visit all j != i within the cutoff; fn(j) is called in cell order.
template <class Fn>
__device__ __forceinline__ void visitNeighbors( const NeighborQuery& q, const CellGrid& g,
                                                const int32_t* __restrict__ cell_start,
                                                const int32_t* __restrict__ list,
                                                const double4* __restrict__ rec, int64_t i, Fn&& fn )
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
                    const double4 r = rec[k];
                    const double dx = r.x - pi0;
                    const double dy = r.y - pi1;
                    const double dz = r.z - pi2;
                    const double dist2 = dx * dx + dy * dy + dz * dz;
                    double cut2 = rc2;
                    if ( q.radius )
                    {
                        const double cut = ri + r.w + q.extend;
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
                                                               const double4* __restrict__ rec,
                                                               int32_t* __restrict__ cnt,
                                                               int64_t* __restrict__ slice_len )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int32_t c = 0;
    if ( p < q.n )
        visitNeighbors( q, g, cell_start, list, rec, p, [&]( int64_t ) { c++; } );
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

// Rows of up to LDSN neighbors are sorted in LDS: every lane appends the
// caller ids of its neighbors to its own LDS column, then places each id at
// its rank (number of smaller ids of the row).  Each lane touches only its own
// column, so no barrier is needed.  Longer rows fall back to an insertion sort
// in the output column itself.  The capped contact lists (at most 32 rows)
// use 8 KB columns, so LDS admits 20 of these one-wave workgroups per CU
// instead of 5.
constexpr int kFillLds = 128;
constexpr int kFillLdsCapped = 32;

// One wavefront (one workgroup) per slice.  CAPPED: the single-pass form of
// the contact lists -- every slice has the same number of rows (the capacity),
// the pass counts the neighbors itself (cnt, the slice maxima in slice_len)
// and raises *overflow when a particle has more than fit.
template <bool CAPPED, int LDSN>
__global__ __launch_bounds__( kSlice ) void kFillNeighbors( NeighborQuery q, CellGrid g,
                                                            const int32_t* __restrict__ cell_start,
                                                            const int32_t* __restrict__ list,
                                                            const double4* __restrict__ rec,
                                                            const int64_t* __restrict__ perm,
                                                            const int64_t* __restrict__ inv,
                                                            const int64_t* __restrict__ slice_ptr,
                                                            int32_t* __restrict__ cnt,
                                                            int32_t* __restrict__ col,
                                                            uint64_t* __restrict__ intact,
                                                            int64_t* __restrict__ slice_len,
                                                            int32_t* __restrict__ overflow )
{
    __shared__ int32_t keys[LDSN * kSlice];
    const int lane = threadIdx.x;
    const int64_t slice = blockIdx.x;
    const int64_t p = slice * kSlice + lane;
    const int64_t base = slice_ptr[slice];
    const int32_t len = static_cast<int32_t>( slice_ptr[slice + 1] - base );
    int32_t* row = col + base * kSlice + lane;
    const bool valid = p < q.n;
    int32_t c = 0;
    bool over = false;
    if ( valid && len <= LDSN )
    {
        int32_t* mine = keys + lane;
        visitNeighbors( q, g, cell_start, list, rec, p, [&]( int64_t j ) {
            if ( c >= len )
            {
                over = true; // CAPPED only: len is otherwise the slice's maximum count
                return;
            }
            mine[c * kSlice] = static_cast<int32_t>( perm[j] );
            c++;
        } );
        for ( int32_t a = 0; a < c; a++ )
        {
            const int32_t key = mine[a * kSlice];
            int32_t r = 0;
            for ( int32_t b = 0; b < c; b++ )
                r += mine[b * kSlice] < key ? 1 : 0;
            row[r * kSlice] = static_cast<int32_t>( inv[key] );
        }
    }
    else if ( valid )
    {
        // Collect caller ids, insertion-sort them, map back to sorted ids.
        visitNeighbors( q, g, cell_start, list, rec, p, [&]( int64_t j ) {
            if ( c >= len )
            {
                over = true;
                return;
            }
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
    int32_t mycnt = 0;
    if constexpr ( CAPPED )
    {
        mycnt = c;
        if ( valid )
            cnt[p] = c;
        int32_t m = c;
        for ( int off = 32; off > 0; off >>= 1 )
            m = max( m, __shfl_xor( m, off ) );
        if ( lane == 0 )
            slice_len[slice] = m;
        if ( __ballot( over ) && lane == 0 )
            *overflow = 1;
    }
    else
        mycnt = valid ? cnt[p] : 0;
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
        if ( valid && slotBond( a, base, k, lane, mycnt ) )
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

// Lattice stencil ids: bond (p, slot) gets the compact id of its integer
// offset (x_j - x_i) / h, looked up in a dense (2K+1)^3 table; 255 marks
// padding.  Any bond whose offset is not a lattice vector of the table sets
// *flag, and the caller then drops the stencil.
__global__ __launch_bounds__( kThreads ) void kStencilIds( PDArgs a, double hx, double hy, double hz, int32_t kx,
                                                           int32_t ky, int32_t kz,
                                                           const int16_t* __restrict__ dense,
                                                           uint8_t* __restrict__ sid, int32_t* __restrict__ flag )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & ( kSlice - 1 );
    const int64_t slice = p / kSlice;
    if ( slice >= a.nslices )
        return;
    const bool valid = p < a.n;
    const int64_t base = a.slice_ptr[slice];
    const int32_t len = static_cast<int32_t>( a.slice_ptr[slice + 1] - base );
    const int32_t mycnt = valid ? a.cnt[p] : 0;
    const int64_t n = a.n;
    const int64_t pc = valid ? p : slice * kSlice;
    const double xi0 = a.x[pc], xi1 = a.x[n + pc], xi2 = a.x[2 * n + pc];
    for ( int32_t q = 0; q < len; q++ )
    {
        uint8_t id = 255;
        if ( q < mycnt )
        {
            const int64_t j = a.col[( base + q ) * kSlice + lane];
            const double gx = a.x[j] - xi0, gy = a.x[n + j] - xi1, gz = a.x[2 * n + j] - xi2;
            const double fa = rint( gx / hx ), fb = rint( gy / hy ), fc = rint( gz / hz );
            const bool on = fabs( gx - fa * hx ) <= 1e-6 * hx && fabs( gy - fb * hy ) <= 1e-6 * hy &&
                            fabs( gz - fc * hz ) <= 1e-6 * hz && fabs( fa ) <= kx && fabs( fb ) <= ky && fabs( fc ) <= kz;
            int16_t c = -1;
            if ( on )
            {
                const int32_t di = ( ( static_cast<int32_t>( fc ) + kz ) * ( 2 * ky + 1 ) + ( static_cast<int32_t>( fb ) + ky ) ) *
                                       ( 2 * kx + 1 ) +
                                   ( static_cast<int32_t>( fa ) + kx );
                c = dense[di];
            }
            if ( c < 0 )
                atomicOr( flag, 1 );
            else
                id = static_cast<uint8_t>( c );
        }
        sid[( base + q ) * kSlice + lane] = id;
    }
}

// ---- Stencil-slot form -------------------------------------------------------
// A full lattice in lattice-row order (device id p = (iz*ny + iy)*nx + ix)
// whose every row lists exactly its in-box stencil entries, in one fixed
// order, is stored with slot q = stencil entry q for every particle: the
// neighbor of slot q is p + off[q] and a per-slot-row word marks which lanes
// have that bond.  ent[q] = (dx, dy, dz, off).
struct LatticeBox
{
    int32_t nx, ny, nz;
};

__device__ __forceinline__ void latticeCoords( int64_t p, const LatticeBox& L, int32_t& ix, int32_t& iy,
                                               int32_t& iz )
{
    const int64_t t = p / L.nx;
    ix = static_cast<int32_t>( p - t * L.nx );
    iy = static_cast<int32_t>( t % L.ny );
    iz = static_cast<int32_t>( t / L.ny );
}

__device__ __forceinline__ bool inBox( const int4 e, int32_t ix, int32_t iy, int32_t iz, const LatticeBox& L )
{
    return static_cast<uint32_t>( ix + e.x ) < static_cast<uint32_t>( L.nx ) &&
           static_cast<uint32_t>( iy + e.y ) < static_cast<uint32_t>( L.ny ) &&
           static_cast<uint32_t>( iz + e.z ) < static_cast<uint32_t>( L.nz );
}

// Every compact row must equal its in-box stencil entries in slot order.
__global__ __launch_bounds__( kThreads ) void kCheckStencilSlots( PDArgs a, LatticeBox L,
                                                                  const int4* __restrict__ ent, int32_t S,
                                                                  int32_t* __restrict__ flag )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p >= a.n )
        return;
    int32_t ix, iy, iz;
    latticeCoords( p, L, ix, iy, iz );
    const int64_t slice = p / kSlice;
    const int lane = static_cast<int>( p - slice * kSlice );
    const int64_t base = a.slice_ptr[slice];
    const int32_t cnt = a.cnt[p];
    int32_t q = 0;
    bool bad = false;
    for ( int32_t e = 0; e < S && !bad; e++ )
    {
        const int4 t = ent[e];
        if ( inBox( t, ix, iy, iz, L ) )
        {
            if ( q >= cnt || a.col[( base + q ) * kSlice + lane] != p + t.w )
                bad = true;
            q++;
        }
    }
    if ( bad || q != cnt )
        *flag = 1;
}

// Fill the slot-form arrays: col (neighbor id or self), exist and intact words.
__global__ __launch_bounds__( kThreads ) void kFillStencilSlots( int64_t n, int64_t nslices, LatticeBox L,
                                                                 const int4* __restrict__ ent, int32_t S,
                                                                 int32_t* __restrict__ col,
                                                                 uint64_t* __restrict__ exist,
                                                                 uint64_t* __restrict__ intact )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t slice = p / kSlice;
    if ( slice >= nslices )
        return; // whole wavefront
    const int lane = threadIdx.x & ( kSlice - 1 );
    const bool valid = p < n;
    int32_t ix = 0, iy = 0, iz = 0;
    if ( valid )
        latticeCoords( p, L, ix, iy, iz );
    const int32_t self = static_cast<int32_t>( valid ? p : slice * kSlice );
    for ( int32_t e = 0; e < S; e++ )
    {
        const int4 t = ent[e];
        const bool ex = valid && inBox( t, ix, iy, iz, L );
        col[( slice * S + e ) * kSlice + lane] = ex ? static_cast<int32_t>( p + t.w ) : self;
        const uint64_t m = __ballot( ex );
        if ( lane == 0 )
        {
            exist[slice * S + e] = m;
            intact[slice * S + e] = m;
        }
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

void launchInvertPerm( const int64_t* perm_dev, int64_t n, int64_t* inv_dev, hipStream_t st )
{
    if ( n == 0 )
        return;
    kInvertPerm<<<gridFor( n ), kThreads, 0, st>>>( perm_dev, n, inv_dev );
    PDK_HIP( hipGetLastError() );
}

bool latticeOrder( const double* x, int64_t n, std::vector<int64_t>& perm, double h_out[3], int64_t dims_out[3],
                   double* dev_out )
{
    if ( n < 2 )
        return false;
    double lo[3], h[3];
    int64_t nk[3] = { 1, 1, 1 };
    for ( int d = 0; d < 3; d++ )
    {
        lo[d] = x[d];
        h[d] = 0.0;
        for ( int64_t p = 0; p < n; p++ )
            lo[d] = std::min( lo[d], x[3 * p + d] );
        double best = 0.0;
        for ( int64_t p = 1; p < n; p++ )
        {
            const double dd = std::fabs( x[3 * p + d] - x[3 * ( p - 1 ) + d] );
            if ( dd > 0.0 && ( best == 0.0 || dd < best ) )
                best = dd;
        }
        h[d] = best > 0.0 ? best : 1.0;
    }
    std::vector<int64_t> k( 3 * n );
    for ( int64_t p = 0; p < n; p++ )
        for ( int d = 0; d < 3; d++ )
        {
            const double q = ( x[3 * p + d] - lo[d] ) / h[d];
            const double r = std::nearbyint( q );
            if ( !( std::fabs( q - r ) <= 1e-6 ) || r > 2e6 )
                return false;
            k[3 * p + d] = static_cast<int64_t>( r );
            nk[d] = std::max( nk[d], k[3 * p + d] + 1 );
        }
    // spacing from the extent (averages out the rounding of single steps)
    for ( int d = 0; d < 3; d++ )
    {
        double hi = lo[d];
        for ( int64_t p = 0; p < n; p++ )
            hi = std::max( hi, x[3 * p + d] );
        h_out[d] = nk[d] > 1 ? ( hi - lo[d] ) / static_cast<double>( nk[d] - 1 ) : h[d];
        dims_out[d] = nk[d];
    }
    double dev = 0.0;
    for ( int d = 0; d < 3; d++ )
    {
        double e = 0.0, mag = 0.0;
        for ( int64_t p = 0; p < n; p++ )
        {
            const double site = lo[d] + static_cast<double>( k[3 * p + d] ) * h_out[d];
            e = std::max( e, std::fabs( x[3 * p + d] - site ) );
            mag = std::max( mag, std::fabs( x[3 * p + d] ) );
        }
        // + 4 ulp of the largest coordinate: the rounding of site and difference
        dev = std::max( dev, ( e + 4.0 * std::ldexp( mag, -52 ) ) / h_out[d] );
    }
    if ( dev_out )
        *dev_out = dev;
    std::vector<uint64_t> key( n );
    bool increasing = true;
    for ( int64_t p = 0; p < n; p++ )
    {
        key[p] = static_cast<uint64_t>( ( k[3 * p + 2] * nk[1] + k[3 * p + 1] ) * nk[0] + k[3 * p] );
        if ( p > 0 && key[p] <= key[p - 1] )
            increasing = false;
    }
    perm.resize( n );
    for ( int64_t p = 0; p < n; p++ )
        perm[p] = p;
    if ( !increasing )
    {
        std::stable_sort( perm.begin(), perm.end(), [&]( int64_t a, int64_t b ) { return key[a] < key[b]; } );
        for ( int64_t p = 1; p < n; p++ )
            if ( key[perm[p]] == key[perm[p - 1]] )
                return false; // two particles on one lattice site
    }
    return true;
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

// slice_ptr[s] = s * cap for s = 0 .. nslices (the capped single-pass lists)
__global__ void kSliceRamp( int64_t* __restrict__ slice_ptr, int64_t nslices, int32_t cap )
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( s <= nslices )
        slice_ptr[s] = s * cap;
}

int64_t buildSlicedCsr( const NeighborQuery& q, const int64_t* perm, const int64_t* inv,
                        int64_t* slice_ptr, int32_t* cnt, DevBuf& col, DevBuf& intact,
                        int64_t& col_capacity_rows, BuildScratch& s, MemAccount* acct,
                        hipStream_t st, int32_t* max_cnt_out, int32_t cap_hint )
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
        double rmax = q.rmax;
        if ( !( rmax > 0 ) )
        {
            std::vector<double> rh( n );
            PDK_HIP( hipMemcpyAsync( rh.data(), q.radius, n * sizeof( double ), hipMemcpyDeviceToHost, st ) );
            PDK_HIP( hipStreamSynchronize( st ) );
            for ( double r : rh )
                rmax = std::max( rmax, r );
        }
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
    ensure<double4>( s.cell_rec, n, acct );
    kCellPack<<<gridFor( n ), kThreads, 0, st>>>( q, s.cell_list.as<int32_t>(), s.cell_rec.as<double4>() );
    PDK_HIP( hipGetLastError() );
    // Single pass into cap_hint rows per slice (contact lists, whose counts
    // change little between rebuilds); a particle with more neighbors than
    // fit sends the build to the counting path below.
    ensure<int64_t>( s.slice_len, nslices + 1, acct );
    if ( cap_hint > 0 && cap_hint <= kFillLdsCapped )
    {
        const int64_t rows = nslices * cap_hint;
        if ( rows > col_capacity_rows || !col )
        {
            col.alloc( static_cast<size_t>( rows ) * kSlice * sizeof( int32_t ), acct );
            intact.alloc( static_cast<size_t>( rows ) * sizeof( uint64_t ), acct );
            col_capacity_rows = rows;
        }
        ensure<int32_t>( s.flag, 1, acct );
        PDK_HIP( hipMemsetAsync( s.flag.as<int32_t>(), 0, sizeof( int32_t ), st ) );
        kSliceRamp<<<static_cast<unsigned>( ( nslices + 1 + kThreads - 1 ) / kThreads ), kThreads, 0, st>>>(
            slice_ptr, nslices, cap_hint );
        PDK_HIP( hipGetLastError() );
        kFillNeighbors<true, kFillLdsCapped><<<static_cast<unsigned>( nslices ), kSlice, 0, st>>>(
            q, g, s.cell_start.as<int32_t>(), s.cell_list.as<int32_t>(), s.cell_rec.as<double4>(), perm, inv,
            slice_ptr, cnt, col.as<int32_t>(), intact.as<uint64_t>(), s.slice_len.as<int64_t>(),
            s.flag.as<int32_t>() );
        PDK_HIP( hipGetLastError() );
        int32_t over = 0;
        PDK_HIP( hipMemcpyAsync( &over, s.flag.as<int32_t>(), sizeof( int32_t ), hipMemcpyDeviceToHost, st ) );
        PDK_HIP( hipStreamSynchronize( st ) );
        if ( !over )
        {
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
            return rows;
        }
    }
    // Counts and slot rows per slice.
    PDK_HIP( hipMemsetAsync( s.slice_len.as<int64_t>(), 0, ( nslices + 1 ) * sizeof( int64_t ), st ) );
    const unsigned gslices = static_cast<unsigned>( ( nslices * kSlice + kThreads - 1 ) / kThreads );
    kCountNeighbors<<<gslices, kThreads, 0, st>>>( q, g, s.cell_start.as<int32_t>(), s.cell_list.as<int32_t>(),
                                                    s.cell_rec.as<double4>(), cnt, s.slice_len.as<int64_t>() );
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
    kFillNeighbors<false, kFillLds><<<static_cast<unsigned>( nslices ), kSlice, 0, st>>>(
        q, g, s.cell_start.as<int32_t>(), s.cell_list.as<int32_t>(), s.cell_rec.as<double4>(), perm, inv, slice_ptr,
        cnt, col.as<int32_t>(), intact.as<uint64_t>(), nullptr, nullptr );
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

bool buildStencil( const PDArgs& a, const double h[3], double rc, int64_t total_rows, DevBuf& sid, DevBuf& table,
                   DevBuf& xi2d, DevBuf& xiv, BuildScratch& s, MemAccount* acct, hipStream_t st, int32_t* nsten_out )
{
    int32_t K[3];
    for ( int d = 0; d < 3; d++ )
    {
        if ( !( h[d] > 0 ) )
            return false;
        K[d] = static_cast<int32_t>( std::ceil( rc / h[d] ) );
        if ( K[d] > 40 )
            return false;
    }
    const int32_t D[3] = { 2 * K[0] + 1, 2 * K[1] + 1, 2 * K[2] + 1 };
    std::vector<int16_t> dense( static_cast<size_t>( D[0] ) * D[1] * D[2], -1 );
    std::vector<float4> tab;
    std::vector<double> tab2;
    std::vector<double> tabv( 4 * 256, 0.0 ); // xi_x, |xi|^2, xi_y, xi_z in FP64 per entry
    const double rc2 = rc * rc * ( 1.0 + 1e-9 );
    for ( int32_t c = -K[2]; c <= K[2]; c++ )
        for ( int32_t b = -K[1]; b <= K[1]; b++ )
            for ( int32_t aa = -K[0]; aa <= K[0]; aa++ )
            {
                const double gx = aa * h[0], gy = b * h[1], gz = c * h[2];
                const double r2 = gx * gx + gy * gy + gz * gz;
                if ( r2 == 0.0 || r2 > rc2 )
                    continue;
                if ( tab.size() >= 255 )
                    return false;
                dense[( ( c + K[2] ) * D[1] + ( b + K[1] ) ) * D[0] + ( aa + K[0] )] = static_cast<int16_t>( tab.size() );
                double* v = &tabv[4 * tab.size()];
                v[kXiX] = gx;
                v[kXi2] = r2;
                v[kXiY] = gy;
                v[kXiZ] = gz;
                tab.push_back( make_float4( static_cast<float>( gx ), static_cast<float>( gy ), static_cast<float>( gz ),
                                            static_cast<float>( std::sqrt( r2 ) ) ) );
                tab2.push_back( r2 );
            }
    while ( tab.size() < 256 )
    {
        tabv[4 * tab.size() + kXi2] = 1.0;
        tab.push_back( make_float4( 0.f, 0.f, 0.f, 1.f ) );
        tab2.push_back( 1.0 );
    }
    DevBuf ddense;
    ddense.alloc( dense.size() * sizeof( int16_t ), acct );
    ensure<int32_t>( s.flag, 1, acct );
    PDK_HIP( hipMemcpyAsync( ddense.as<int16_t>(), dense.data(), dense.size() * sizeof( int16_t ), hipMemcpyHostToDevice, st ) );
    PDK_HIP( hipMemsetAsync( s.flag.as<int32_t>(), 0, sizeof( int32_t ), st ) );
    sid.alloc( static_cast<size_t>( std::max<int64_t>( total_rows, 1 ) ) * kSlice, acct );
    table.alloc( 256 * sizeof( float4 ), acct );
    PDK_HIP( hipMemcpyAsync( table.as<float4>(), tab.data(), 256 * sizeof( float4 ), hipMemcpyHostToDevice, st ) );
    xi2d.alloc( 256 * sizeof( double ), acct );
    PDK_HIP( hipMemcpyAsync( xi2d.as<double>(), tab2.data(), 256 * sizeof( double ), hipMemcpyHostToDevice, st ) );
    xiv.alloc( tabv.size() * sizeof( double ), acct );
    PDK_HIP( hipMemcpyAsync( xiv.as<double>(), tabv.data(), tabv.size() * sizeof( double ), hipMemcpyHostToDevice, st ) );
    const unsigned g = static_cast<unsigned>( ( a.nslices * kSlice + kThreads - 1 ) / kThreads );
    kStencilIds<<<g, kThreads, 0, st>>>( a, h[0], h[1], h[2], K[0], K[1], K[2], ddense.as<int16_t>(), sid.as<uint8_t>(),
                                         s.flag.as<int32_t>() );
    PDK_HIP( hipGetLastError() );
    int32_t flag = 0;
    PDK_HIP( hipMemcpyAsync( &flag, s.flag.as<int32_t>(), sizeof( int32_t ), hipMemcpyDeviceToHost, st ) );
    PDK_HIP( hipStreamSynchronize( st ) );
    if ( flag )
    {
        sid.release();
        table.release();
        xi2d.release();
        xiv.release();
        return false;
    }
    if ( nsten_out )
    {
        int32_t cnt = 0;
        for ( const float4& t : tab )
            cnt += ( t.w != 1.f || t.x != 0.f || t.y != 0.f || t.z != 0.f ) ? 1 : 0;
        *nsten_out = cnt;
    }
    return true;
}

namespace
{
struct SlotEntry
{
    int32_t a, b, c;
    int64_t off;
    double r2;
};

// Stencil entries of a full lattice in slot order.  margin: relative band
// around the cutoff (in squared distance) inside which no lattice vector may
// lie, since there the rounding of the actual positions decides.
bool slotStencil( int64_t n, const int64_t dims[3], const double h[3], double rc,
                  const std::vector<int64_t>& perm_host, double margin, std::vector<SlotEntry>& es,
                  LatticeBox& L, bool& identity )
{
    if ( dims[0] * dims[1] * dims[2] != n || static_cast<int64_t>( perm_host.size() ) != n )
        return false;
    // Caller order equal to the lattice order: slots sorted by id offset are
    // sorted by caller id for every particle.  Otherwise the slot order is
    // taken from an interior particle (which must exist) and verified.
    identity = true;
    for ( int64_t p = 0; p < n && identity; p++ )
        identity = perm_host[p] == p;
    int32_t K[3];
    for ( int d = 0; d < 3; d++ )
    {
        if ( !( h[d] > 0 ) )
            return false;
        // largest integer offset along d that can be within the cutoff
        K[d] = static_cast<int32_t>( std::floor( rc / h[d] * ( 1.0 + margin ) ) );
        if ( K[d] > 40 || ( !identity && dims[d] <= 2 * K[d] ) )
            return false;
    }
    L = LatticeBox{ static_cast<int32_t>( dims[0] ), static_cast<int32_t>( dims[1] ), static_cast<int32_t>( dims[2] ) };
    es.clear();
    const double rc2 = rc * rc;
    for ( int32_t c = -K[2]; c <= K[2]; c++ )
        for ( int32_t b = -K[1]; b <= K[1]; b++ )
            for ( int32_t aa = -K[0]; aa <= K[0]; aa++ )
            {
                const double gx = aa * h[0], gy = b * h[1], gz = c * h[2];
                const double r2 = gx * gx + gy * gy + gz * gz;
                if ( r2 == 0.0 )
                    continue;
                if ( std::fabs( r2 - rc2 ) <= margin * rc2 )
                    return false;
                if ( r2 > rc2 )
                    continue;
                // offsets that no particle of the box can have
                if ( std::abs( aa ) >= dims[0] || std::abs( b ) >= dims[1] || std::abs( c ) >= dims[2] )
                    continue;
                es.push_back( { aa, b, c, ( static_cast<int64_t>( c ) * L.ny + b ) * L.nx + aa, r2 } );
            }
    const int32_t S = static_cast<int32_t>( es.size() );
    if ( S == 0 || S > 255 )
        return false;
    if ( identity )
        std::stable_sort( es.begin(), es.end(), []( const SlotEntry& u, const SlotEntry& v ) { return u.off < v.off; } );
    else
    {
        // ascending caller id of an interior particle's neighbors
        const int64_t p0 = ( static_cast<int64_t>( K[2] ) * L.ny + K[1] ) * L.nx + K[0];
        std::stable_sort( es.begin(), es.end(), [&]( const SlotEntry& u, const SlotEntry& v ) {
            return perm_host[p0 + u.off] < perm_host[p0 + v.off];
        } );
    }
    return true;
}

std::vector<int4> slotEntries( const std::vector<SlotEntry>& es )
{
    std::vector<int4> ent( es.size() );
    for ( size_t e = 0; e < es.size(); e++ )
        ent[e] = make_int4( es[e].a, es[e].b, es[e].c, static_cast<int32_t>( es[e].off ) );
    return ent;
}

// Fill the slot arrays, the row table, offsets, the FP32 stencil table and
// the FP64 |xi|^2 of each entry, and move them into place.
void finishSlots( int64_t n, int64_t nslices, const LatticeBox& L, const std::vector<SlotEntry>& es,
                  const int4* dent, const double h[3], DevBuf& col, DevBuf& intact, int64_t* slice_ptr_dev,
                  int64_t& total_rows, int64_t& col_capacity_rows, SlotForm& out, MemAccount* acct,
                  hipStream_t st )
{
    const int32_t S = static_cast<int32_t>( es.size() );
    const int64_t rows = nslices * S;
    DevBuf ncol, nintact, nexist, dsoff, dtab;
    ncol.alloc( static_cast<size_t>( rows ) * kSlice * sizeof( int32_t ), acct );
    nintact.alloc( static_cast<size_t>( rows ) * sizeof( uint64_t ), acct );
    nexist.alloc( static_cast<size_t>( rows ) * sizeof( uint64_t ), acct );
    kFillStencilSlots<<<gridFor( nslices * kSlice ), kThreads, 0, st>>>( n, nslices, L, dent, S, ncol.as<int32_t>(),
                                                                         nexist.as<uint64_t>(), nintact.as<uint64_t>() );
    PDK_HIP( hipGetLastError() );
    std::vector<int64_t> sp( nslices + 1 );
    for ( int64_t k = 0; k <= nslices; k++ )
        sp[k] = k * S;
    PDK_HIP( hipMemcpyAsync( slice_ptr_dev, sp.data(), ( nslices + 1 ) * sizeof( int64_t ), hipMemcpyHostToDevice, st ) );
    // Row windows: maximal runs of slots with one (dy, dz) and consecutive dx.
    std::vector<int4> rtab;
    for ( int32_t e = 0; e < S; e++ )
    {
        if ( !rtab.empty() )
        {
            int4& r = rtab.back();
            const int32_t dx0 = ( r.w & 0xff ) - 128, dy = ( ( r.w >> 8 ) & 0xff ) - 128,
                          dz = ( ( r.w >> 16 ) & 0xff ) - 128;
            if ( es[e].b == dy && es[e].c == dz && es[e].a == dx0 + r.y && r.y < kRowMaxSlots )
            {
                r.y++;
                continue;
            }
        }
        const int32_t packed = ( es[e].a + 128 ) | ( ( es[e].b + 128 ) << 8 ) | ( ( es[e].c + 128 ) << 16 );
        rtab.push_back( make_int4( e, 1, static_cast<int32_t>( es[e].off ), packed ) );
    }
    DevBuf drows, drstart;
    int32_t nrows = 0;
    // the row-window passes address their windows with 32-bit byte offsets
    // checked against 32-bit buffer sizes: n < 2^28
    if ( static_cast<int>( rtab.size() ) <= kMaxRows && n < ( int64_t( 1 ) << 28 ) )
    {
        nrows = static_cast<int32_t>( rtab.size() );
        drows.alloc( nrows * sizeof( int4 ), acct );
        PDK_HIP( hipMemcpyAsync( drows.as<int4>(), rtab.data(), nrows * sizeof( int4 ), hipMemcpyHostToDevice, st ) );
        std::vector<uint32_t> rs( 9, 0u );
        for ( const int4& r : rtab )
            rs[r.x >> 5] |= 1u << ( r.x & 31 );
        drstart.alloc( 9 * sizeof( uint32_t ), acct );
        PDK_HIP( hipMemcpyAsync( drstart.as<uint32_t>(), rs.data(), 9 * sizeof( uint32_t ), hipMemcpyHostToDevice, st ) );
    }
    std::vector<int32_t> soff( 256, 0 );
    std::vector<float4> tab( 256, make_float4( 0.f, 0.f, 0.f, 1.f ) );
    std::vector<double> tab2( 256, 1.0 );
    std::vector<double> tabv( 4 * 256, 0.0 ); // xi_x, |xi|^2, xi_y, xi_z in FP64 per entry
    for ( int32_t e = 0; e < S; e++ )
    {
        soff[e] = static_cast<int32_t>( es[e].off );
        const double gx = es[e].a * h[0], gy = es[e].b * h[1], gz = es[e].c * h[2];
        tab[e] = make_float4( static_cast<float>( gx ), static_cast<float>( gy ), static_cast<float>( gz ),
                              static_cast<float>( std::sqrt( es[e].r2 ) ) );
        tab2[e] = gx * gx + gy * gy + gz * gz;
        tabv[4 * e + kXiX] = gx;
        tabv[4 * e + kXiY] = gy;
        tabv[4 * e + kXiZ] = gz;
        tabv[4 * e + kXi2] = tab2[e];
    }
    DevBuf dxiv;
    dxiv.alloc( tabv.size() * sizeof( double ), acct );
    PDK_HIP( hipMemcpyAsync( dxiv.as<double>(), tabv.data(), tabv.size() * sizeof( double ), hipMemcpyHostToDevice,
                             st ) );
    DevBuf dxi2;
    dsoff.alloc( 256 * sizeof( int32_t ), acct );
    dtab.alloc( 256 * sizeof( float4 ), acct );
    dxi2.alloc( 256 * sizeof( double ), acct );
    PDK_HIP( hipMemcpyAsync( dsoff.as<int32_t>(), soff.data(), 256 * sizeof( int32_t ), hipMemcpyHostToDevice, st ) );
    PDK_HIP( hipMemcpyAsync( dtab.as<float4>(), tab.data(), 256 * sizeof( float4 ), hipMemcpyHostToDevice, st ) );
    PDK_HIP( hipMemcpyAsync( dxi2.as<double>(), tab2.data(), 256 * sizeof( double ), hipMemcpyHostToDevice, st ) );
    PDK_HIP( hipStreamSynchronize( st ) );
    col = std::move( ncol );
    intact = std::move( nintact );
    out.exist = std::move( nexist );
    out.soff = std::move( dsoff );
    out.table = std::move( dtab );
    out.xi2d = std::move( dxi2 );
    out.xiv = std::move( dxiv );
    out.rows = std::move( drows );
    out.rowstart = std::move( drstart );
    out.nrows = nrows;
    out.slots = S;
    total_rows = rows;
    col_capacity_rows = rows;
}

// Direct build check: every in-box stencil bond passes the exact FP64 cutoff
// test (the cell search's, CabanaPD_Neighbor.hpp:46,56-59) and, for a caller
// order other than the lattice order, lists its neighbors in ascending
// caller id; also writes the bond count of each particle.
__global__ __launch_bounds__( kThreads ) void kDirectSlotCheck( const double* __restrict__ x, int64_t n, LatticeBox L,
                                                                const int4* __restrict__ ent, int32_t S, double rc2,
                                                                const int64_t* __restrict__ perm, int32_t* __restrict__ cnt,
                                                                int32_t* __restrict__ flag )
{
#pragma clang fp contract( off )
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p >= n )
        return;
    int32_t ix, iy, iz;
    latticeCoords( p, L, ix, iy, iz );
    const double pi0 = x[p], pi1 = x[n + p], pi2 = x[2 * n + p];
    int32_t c = 0;
    int64_t last = -1;
    bool bad = false;
    for ( int32_t e = 0; e < S; e++ )
    {
        const int4 t = ent[e];
        if ( !inBox( t, ix, iy, iz, L ) )
            continue;
        const int64_t j = p + t.w;
        const double dx = x[j] - pi0;
        const double dy = x[n + j] - pi1;
        const double dz = x[2 * n + j] - pi2;
        if ( !( dx * dx + dy * dy + dz * dz <= rc2 ) )
            bad = true;
        if ( perm )
        {
            const int64_t cid = perm[j];
            if ( cid <= last )
                bad = true;
            last = cid;
        }
        c++;
    }
    cnt[p] = c;
    if ( bad )
        *flag = 1;
}
} // namespace

bool buildStencilSlots( const PDArgs& a, const int64_t dims[3], const double h[3], double rc,
                        const std::vector<int64_t>& perm_host, DevBuf& col, DevBuf& intact, int64_t* slice_ptr_dev,
                        int64_t& total_rows, int64_t& col_capacity_rows, SlotForm& out, BuildScratch& s,
                        MemAccount* acct, hipStream_t st )
{
    const int64_t n = a.n;
    std::vector<SlotEntry> es;
    LatticeBox L{};
    bool identity = false;
    if ( !slotStencil( n, dims, h, rc, perm_host, 1e-6, es, L, identity ) )
        return false;
    const int32_t S = static_cast<int32_t>( es.size() );
    const std::vector<int4> ent = slotEntries( es );
    DevBuf dent;
    dent.alloc( S * sizeof( int4 ), acct );
    ensure<int32_t>( s.flag, 1, acct );
    PDK_HIP( hipMemcpyAsync( dent.as<int4>(), ent.data(), S * sizeof( int4 ), hipMemcpyHostToDevice, st ) );
    PDK_HIP( hipMemsetAsync( s.flag.as<int32_t>(), 0, sizeof( int32_t ), st ) );
    kCheckStencilSlots<<<gridFor( n ), kThreads, 0, st>>>( a, L, dent.as<int4>(), S, s.flag.as<int32_t>() );
    PDK_HIP( hipGetLastError() );
    int32_t flag = 0;
    PDK_HIP( hipMemcpyAsync( &flag, s.flag.as<int32_t>(), sizeof( int32_t ), hipMemcpyDeviceToHost, st ) );
    PDK_HIP( hipStreamSynchronize( st ) );
    if ( flag )
        return false;
    finishSlots( n, a.nslices, L, es, dent.as<int4>(), h, col, intact, slice_ptr_dev, total_rows, col_capacity_rows,
                 out, acct, st );
    return true;
}

bool buildStencilSlotsDirect( const double* x_dev, int64_t n, const int64_t dims[3], const double h[3], double rc,
                              const std::vector<int64_t>& perm_host, const int64_t* perm_dev, int32_t* cnt_dev,
                              DevBuf& col, DevBuf& intact, int64_t* slice_ptr_dev, int64_t& total_rows,
                              int64_t& col_capacity_rows, SlotForm& out, BuildScratch& s, MemAccount* acct,
                              hipStream_t st )
{
    // A wider band than the verified build's: positions deviate from the
    // lattice by at most ~1e-6 h per axis, so no pair of particles whose
    // lattice vector lies outside the band can fall on the other side of
    // the cutoff; inside the stencil the exact test is applied anyway.
    std::vector<SlotEntry> es;
    LatticeBox L{};
    bool identity = false;
    if ( !slotStencil( n, dims, h, rc, perm_host, 2e-5, es, L, identity ) )
        return false;
    const int32_t S = static_cast<int32_t>( es.size() );
    const std::vector<int4> ent = slotEntries( es );
    DevBuf dent;
    dent.alloc( S * sizeof( int4 ), acct );
    ensure<int32_t>( s.flag, 1, acct );
    PDK_HIP( hipMemcpyAsync( dent.as<int4>(), ent.data(), S * sizeof( int4 ), hipMemcpyHostToDevice, st ) );
    PDK_HIP( hipMemsetAsync( s.flag.as<int32_t>(), 0, sizeof( int32_t ), st ) );
    kDirectSlotCheck<<<gridFor( n ), kThreads, 0, st>>>( x_dev, n, L, dent.as<int4>(), S, rc * rc,
                                                         identity ? nullptr : perm_dev, cnt_dev, s.flag.as<int32_t>() );
    PDK_HIP( hipGetLastError() );
    int32_t flag = 0;
    PDK_HIP( hipMemcpyAsync( &flag, s.flag.as<int32_t>(), sizeof( int32_t ), hipMemcpyDeviceToHost, st ) );
    PDK_HIP( hipStreamSynchronize( st ) );
    if ( flag )
        return false;
    finishSlots( n, ( n + kSlice - 1 ) / kSlice, L, es, dent.as<int4>(), h, col, intact, slice_ptr_dev, total_rows,
                 col_capacity_rows, out, acct, st );
    return true;
}

void launchPrenotch( const PDArgs& a, int32_t nnotch, const double* notch_dev, hipStream_t st )
{
    const unsigned g = static_cast<unsigned>( ( a.nslices * kSlice + kThreads - 1 ) / kThreads );
    kPrenotch<<<g, kThreads, 0, st>>>( a, nnotch, notch_dev );
    PDK_HIP( hipGetLastError() );
}

} // namespace pdk
