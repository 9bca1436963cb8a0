// pdk_output.hip -- region outputs on the device: region membership from the
// reference positions, reductions of a per-particle quantity over a region
// into a device-resident time series, and gathers of the quantity over the
// members (line profiles).
//
// Reference behaviour restated:
//   Region<RectangularPrism>::inside  CabanaPD_Geometry.hpp:54-82 (inclusive bounds)
//   Region<Cylinder>::inside          CabanaPD_Geometry.hpp:84-135 (inclusive radii and z)
//   Region<Line>::inside              CabanaPD_Geometry.hpp:137-168 (strict half widths)
//   OutputTimeSeries::update          CabanaPD_OutputProfiles.hpp:146-163 (Kokkos::Sum/Max
//                                     reducer joined over the steering vector)
//   createOutputProfile               CabanaPD_OutputProfiles.hpp:26-62
//
// The reductions are two fixed-shape passes (per-block partials, then one
// block), so a series is reproducible run to run; Max and Min are exact.
// Max/Min follow Kokkos's join (replace when strictly greater / smaller,
// starting from the reduction identity -DBL_MAX / +DBL_MAX), so NaN values
// never win.  These kernels stream the member list and one or three fields:
// HBM-bound gathers, a few bytes per member.
#include <cfloat>

#include "pdk_internal.hpp"

namespace pdk
{
namespace
{
constexpr int kOutThreads = 256;

__device__ __forceinline__ bool insideRegion( const RegionParams& R, double x0, double x1, double x2 )
{
    switch ( R.kind )
    {
    case kRegionBox:
        return x0 >= R.low[0] && x0 <= R.high[0] && x1 >= R.low[1] && x1 <= R.high[1] && x2 >= R.low[2] &&
               x2 <= R.high[2];
    case kRegionCylinder:
    {
        const double dx = x0 - R.x_center;
        const double dy = x1 - R.y_center;
        const double rsq = dx * dx + dy * dy;
        return rsq >= R.r_in * R.r_in && rsq <= R.r_out * R.r_out && x2 >= R.low_z && x2 <= R.high_z;
    }
    case kRegionLine:
    {
        const double xs[3] = { x0, x1, x2 };
        const double a = xs[( R.dim + 1 ) % 3];
        const double b = xs[( R.dim + 2 ) % 3];
        return a < R.half[0] && a > -R.half[0] && b < R.half[1] && b > -R.half[1];
    }
    default:
        return true; // every particle
    }
}

// mask[c] for caller id c (device index inv[c]); x is [3][n] in device order.
__global__ void kRegionMask( RegionParams R, const double* __restrict__ x, const int64_t* __restrict__ inv, int64_t n,
                             uint8_t* __restrict__ mask )
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( c >= n )
        return;
    const int64_t p = inv[c];
    mask[c] = insideRegion( R, x[p], x[n + p], x[2 * n + p] ) ? 1 : 0;
}

__device__ __forceinline__ double quantityAt( const QuantityArgs& Q, int64_t p )
{
    double v = Q.src[p];
    if ( Q.gate )
        v = Q.gate[p] > Q.threshold ? v : 0.0;
    if ( Q.vol )
        v *= Q.vol[p];
    return v;
}

__device__ __forceinline__ double identityOf( int32_t red )
{
    return red == kReduceMax ? -DBL_MAX : ( red == kReduceMin ? DBL_MAX : 0.0 );
}

__device__ __forceinline__ double join( int32_t red, double acc, double v )
{
    if ( red == kReduceMax )
        return v > acc ? v : acc;
    if ( red == kReduceMin )
        return v < acc ? v : acc;
    return acc + v;
}

// Fixed-order tree over the block's values (blockDim.x == kOutThreads).
__device__ double blockJoin( int32_t red, double v )
{
    __shared__ double lds[kOutThreads];
    lds[threadIdx.x] = v;
    __syncthreads();
    for ( int w = kOutThreads / 2; w > 0; w >>= 1 )
    {
        if ( threadIdx.x < w )
            lds[threadIdx.x] = join( red, lds[threadIdx.x], lds[threadIdx.x + w] );
        __syncthreads();
    }
    return lds[0];
}

__global__ void __launch_bounds__( kOutThreads )
    kSeriesPartial( const int64_t* __restrict__ idx, int64_t count, QuantityArgs Q, int32_t red,
                    double* __restrict__ partial )
{
    double acc = identityOf( red );
    const int64_t stride = (int64_t)gridDim.x * kOutThreads;
    for ( int64_t k = blockIdx.x * (int64_t)kOutThreads + threadIdx.x; k < count; k += stride )
        acc = join( red, acc, quantityAt( Q, idx[k] ) );
    acc = blockJoin( red, acc );
    if ( threadIdx.x == 0 )
        partial[blockIdx.x] = acc;
}

__global__ void __launch_bounds__( kOutThreads )
    kSeriesFinal( const double* __restrict__ partial, int32_t nparts, int32_t red, double* __restrict__ out )
{
    double acc = identityOf( red );
    for ( int32_t k = threadIdx.x; k < nparts; k += kOutThreads )
        acc = join( red, acc, partial[k] );
    acc = blockJoin( red, acc );
    if ( threadIdx.x == 0 )
        *out = acc;
}

__global__ void kGatherQuantity( const int64_t* __restrict__ idx, int64_t count, QuantityArgs Q,
                                 double* __restrict__ out )
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( k < count )
        out[k] = quantityAt( Q, idx[k] );
}

__global__ void kScatterRows( const double* __restrict__ src, const int64_t* __restrict__ perm,
                              double* __restrict__ dst, int64_t n, int32_t ncomp )
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( p >= n )
        return;
    const int64_t c = perm[p];
    for ( int32_t k = 0; k < ncomp; k++ )
        dst[ncomp * c + k] = src[k * n + p];
}

unsigned gridOf( int64_t n )
{
    return static_cast<unsigned>( ( n + kOutThreads - 1 ) / kOutThreads );
}
} // namespace

void launchRegionMask( const RegionParams& R, const double* x_dev, const int64_t* inv_dev, int64_t n,
                       uint8_t* mask_dev, hipStream_t st )
{
    if ( n == 0 )
        return;
    kRegionMask<<<gridOf( n ), kOutThreads, 0, st>>>( R, x_dev, inv_dev, n, mask_dev );
    PDK_HIP( hipGetLastError() );
}

int32_t seriesPartials( int64_t count )
{
    const int64_t b = ( count + kOutThreads - 1 ) / kOutThreads;
    return static_cast<int32_t>( b < 1 ? 1 : ( b > kSeriesMaxPartials ? kSeriesMaxPartials : b ) );
}

void launchSeriesReduce( const int64_t* idx_dev, int64_t count, const QuantityArgs& Q, int32_t red,
                         double* partial_dev, double* out_dev, hipStream_t st )
{
    const int32_t nparts = seriesPartials( count );
    kSeriesPartial<<<nparts, kOutThreads, 0, st>>>( idx_dev, count, Q, red, partial_dev );
    PDK_HIP( hipGetLastError() );
    kSeriesFinal<<<1, kOutThreads, 0, st>>>( partial_dev, nparts, red, out_dev );
    PDK_HIP( hipGetLastError() );
}

void launchScatterRows( const double* src_soa, const int64_t* perm, double* dst_aos, int64_t n, int32_t ncomp,
                        hipStream_t st )
{
    if ( n == 0 )
        return;
    kScatterRows<<<gridOf( n ), kOutThreads, 0, st>>>( src_soa, perm, dst_aos, n, ncomp );
    PDK_HIP( hipGetLastError() );
}

void launchGatherQuantity( const int64_t* idx_dev, int64_t count, const QuantityArgs& Q, double* out_dev,
                           hipStream_t st )
{
    if ( count == 0 )
        return;
    kGatherQuantity<<<gridOf( count ), kOutThreads, 0, st>>>( idx_dev, count, Q, out_dev );
    PDK_HIP( hipGetLastError() );
}

} // namespace pdk
