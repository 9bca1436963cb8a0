// pdk_internal.hpp -- host-side internals shared by the engine's translation
// units (device buffers, error reporting, kernel launch wrappers).
#ifndef PDK_INTERNAL_HPP
#define PDK_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <mutex>
#include <string>
#include <vector>

#include "pdk_common.hpp"

namespace pdk
{
// Thrown inside the library, turned into a status code at the C boundary.
struct Error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

#define PDK_HIP( call )                                                                        \
    do                                                                                         \
    {                                                                                          \
        hipError_t pdk_e_ = ( call );                                                          \
        if ( pdk_e_ != hipSuccess )                                                            \
            throw ::pdk::Error( std::string( #call " failed: " ) + hipGetErrorString( pdk_e_ ) ); \
    } while ( 0 )

// Accounting hook: every device allocation of a spec goes through its pool so
// the context can enforce the HBM budget.
// HBM accounting of one context.  Specimens may be set up on one host thread
// while another drives the batch engine, so updates are serialized.
struct MemAccount
{
    uint64_t budget = 0;
    uint64_t in_use = 0;
    std::mutex mu;

    void reserve( uint64_t bytes )
    {
        std::lock_guard<std::mutex> g( mu );
        if ( budget && in_use + bytes > budget )
            throw Error( "HBM budget exceeded: need " + std::to_string( bytes ) + " B more, " +
                         std::to_string( budget - in_use ) + " B left" );
        in_use += bytes;
    }
    void unreserve( uint64_t bytes )
    {
        std::lock_guard<std::mutex> g( mu );
        in_use -= bytes;
    }
    uint64_t used()
    {
        std::lock_guard<std::mutex> g( mu );
        return in_use;
    }
};

// Owning device buffer.
class DevBuf
{
  public:
    DevBuf() = default;
    DevBuf( const DevBuf& ) = delete;
    DevBuf& operator=( const DevBuf& ) = delete;
    DevBuf( DevBuf&& o ) noexcept { swap( o ); }
    DevBuf& operator=( DevBuf&& o ) noexcept
    {
        release();
        swap( o );
        return *this;
    }
    ~DevBuf() { release(); }

    void alloc( size_t bytes, MemAccount* acct )
    {
        release();
        if ( bytes == 0 )
            return;
        if ( acct )
            acct->reserve( bytes );
        const hipError_t e = hipMalloc( &ptr_, bytes );
        if ( e != hipSuccess )
        {
            ptr_ = nullptr;
            if ( acct )
                acct->unreserve( bytes );
            PDK_HIP( e );
        }
        bytes_ = bytes;
        acct_ = acct;
    }
    void release()
    {
        if ( ptr_ )
        {
            (void)hipFree( ptr_ );
            if ( acct_ )
                acct_->unreserve( bytes_ );
        }
        ptr_ = nullptr;
        bytes_ = 0;
        acct_ = nullptr;
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>( ptr_ );
    }
    size_t bytes() const { return bytes_; }
    explicit operator bool() const { return ptr_ != nullptr; }

  private:
    void swap( DevBuf& o )
    {
        std::swap( ptr_, o.ptr_ );
        std::swap( bytes_, o.bytes_ );
        std::swap( acct_, o.acct_ );
    }
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
    MemAccount* acct_ = nullptr;
};

// ---- bond build (pdk_build.hip) ---------------------------------------------
struct BuildScratch
{
    DevBuf keys, keys2, vals, vals2, tmp, cell, cell_start, cell_fill, cell_list, cell_rec, slice_len, bbox,
        flag;
};

// Morton order of the caller's positions: perm[p] = caller id of sorted p.
void launchMortonOrder( const double* x_aos_dev, int64_t n, const double lo[3], double inv_h,
                        int64_t* perm_dev, int64_t* inv_dev, BuildScratch& s, MemAccount* acct,
                        hipStream_t st );
// Lattice detection: true when every position is lo + k h (k integer per
// axis, one particle per site); perm then lists particles in lattice index
// order with x fastest.  dev_out: the largest distance of a position from
// its site lo + k h along an axis, as a fraction of that axis' spacing, plus
// the rounding of the site (stencil bond vectors differ from x_j - x_i by at
// most twice that per axis).
bool latticeOrder( const double* x_aos_host, int64_t n, std::vector<int64_t>& perm, double h_out[3],
                   int64_t dims_out[3], double* dev_out );
// Stencil ids of a lattice specimen's bonds (mixed form); false when some
// bond is not a lattice vector or the stencil has more than 255 entries.
bool buildStencil( const PDArgs& a, const double h[3], double rc, int64_t total_rows, DevBuf& sid, DevBuf& table,
                   DevBuf& xi2d, DevBuf& xiv, BuildScratch& s, MemAccount* acct, hipStream_t st, int32_t* nsten_out );
// Stencil-slot form of a full lattice (dims[0]*dims[1]*dims[2] == n, lattice
// rows): verifies that every compact row equals its in-box stencil entries in
// one fixed order (the caller-id order of an interior particle's row), then
// rebuilds col / intact / slice_ptr with slot q = stencil entry q and fills
// exist, the slot offsets soff[256], the FP32 stencil table[256] in slot
// order and the row-window table.  False (arrays untouched) when the
// specimen does not qualify.
struct SlotForm
{
    DevBuf exist, soff, table, xi2d, xiv, rows, rowstart, aux, rec;
    int32_t slots = 0;
    int32_t nrows = 0; // row windows (0: the stencil does not split into <= kMaxRows rows)
};
bool buildStencilSlots( const PDArgs& a, const int64_t dims[3], const double h[3], double rc,
                        const std::vector<int64_t>& perm_host, DevBuf& col, DevBuf& intact, int64_t* slice_ptr_dev,
                        int64_t& total_rows, int64_t& col_capacity_rows, SlotForm& out, BuildScratch& s,
                        MemAccount* acct, hipStream_t st );
// The same form built without the cell search: every in-box stencil bond is
// checked with the exact cutoff test; fills cnt.  False (nothing written but
// cnt) when the lattice does not qualify.
bool buildStencilSlotsDirect( const double* x_dev, int64_t n, const int64_t dims[3], const double h[3], double rc,
                              const std::vector<int64_t>& perm_host, const int64_t* perm_dev, int32_t* cnt_dev,
                              DevBuf& col, DevBuf& intact, int64_t* slice_ptr_dev, int64_t& total_rows,
                              int64_t& col_capacity_rows, SlotForm& out, BuildScratch& s, MemAccount* acct,
                              hipStream_t st );
void launchInvertPerm( const int64_t* perm_dev, int64_t n, int64_t* inv_dev, hipStream_t st );
void launchGatherAos3( const double* src_aos, const int64_t* perm, double* dst_soa, int64_t n,
                       hipStream_t st );
void launchScatterAos3( const double* src_soa, const int64_t* perm, double* dst_aos, int64_t n,
                        hipStream_t st );
void launchGatherScalar( const double* src, const int64_t* perm, double* dst, int64_t n,
                         hipStream_t st );
void launchScatterScalar( const double* src, const int64_t* perm, double* dst, int64_t n,
                          hipStream_t st );
void launchGatherFlags( const int32_t* src, const int64_t* perm, uint8_t* dst, int64_t n,
                        hipStream_t st );

struct NeighborQuery
{
    const double* x;      // sorted SoA positions
    const double* u;      // if non-null, positions are x + u
    const double* radius; // polydisperse contact: pair cutoff R_i + R_j + extend
    double rc;            // cutoff (uniform)
    double extend;        // polydisperse: added to R_i + R_j
    int64_t n;
    double rmax = 0.0;    // polydisperse: largest radius if known (else read back from the device)
};

// Builds the sliced CSR; (re)allocates col/intact when capacity is short.
// Returns total slot rows.
// cap_hint > 0: try one pass into cap_hint rows per slice first (contact lists).
int64_t buildSlicedCsr( const NeighborQuery& q, const int64_t* perm, const int64_t* inv,
                        int64_t* slice_ptr, int32_t* cnt, DevBuf& col, DevBuf& intact,
                        int64_t& col_capacity_rows, BuildScratch& s, MemAccount* acct,
                        hipStream_t st, int32_t* max_cnt_out, int32_t cap_hint = 0 );

// Prenotch<N>::create over the sliced CSR.
void launchPrenotch( const PDArgs& a, int32_t nnotch, const double* notch_dev, hipStream_t st );

// ---- PD passes (pdk_force.hip) -----------------------------------------------
void launchForce( const PDArgs* args_dev, const PDArgs& a, int32_t epilogue, int32_t parity,
                  hipStream_t st );
void launchLpsState( const PDArgs* args_dev, const PDArgs& a, bool weighted_volume, bool dilatation,
                     int32_t parity, hipStream_t st );
void launchEnergy( const PDArgs* args_dev, const PDArgs& a, int32_t parity, double* partial_dev,
                   int64_t nblocks, hipStream_t st );
int64_t energyBlocks( const PDArgs& a );
// Per stencil entry (1/|xi|, sqrt(3) bandf/|xi|, |xi|^2, 0) for the mixed row-window passes.
void launchStencilAux( const float4* table, int32_t S, int32_t influence, float4* aux, hipStream_t st );
// The row-window record of each stencil entry (kRec* layout) from the FP32
// table, its aux table and the FP64 offsets.
void launchStencilRec( const float4* table, const float4* aux, const double* xiv, int32_t S, double* rec,
                       hipStream_t st );
// Stress tensor pass into a.stress ([9][n], device order).
void launchStress( const PDArgs* args_dev, const PDArgs& a, int32_t parity, hipStream_t st );
const char* stressKernelName( const PDArgs& a );
// Scatter a [ncomp][n] device-order array to a caller-order AoS [n][ncomp].
void launchScatterRows( const double* src_soa, const int64_t* perm, double* dst_aos, int64_t n, int32_t ncomp,
                        hipStream_t st );
void launchInitialHalfStep( const PDArgs* args_dev, const PDArgs& a, int32_t parity, hipStream_t st );
// Contact list of a PD specimen (pdk_set_contact) and its NormalRepulsion
// constants, passed by value to the contact pass.
struct PdContactArgs
{
    const int64_t* slice_ptr;
    const int32_t* cnt;
    const int32_t* col;
    double radius, delta, c;
};
// Adds the contact forces to the stored PD force, then the epilogue (force
// terms and velocity-Verlet halves) of mode epi.
void launchPdContact( const PDArgs* args_dev, const PDArgs& a, const PdContactArgs& ca, int32_t epi, int32_t parity,
                      hipStream_t st );
// max_p |U[p] - u_neigh[p]| into *out_dev (partial holds one double per block of 256).
void launchMaxDisplacement( const double* U, const double* u_neigh, int64_t n, double* partial, double* out_dev,
                            hipStream_t st );
int64_t maxDisplacementBlocks( int64_t n );
void launchApplyForceTerms( const PDArgs* args_dev, const PDArgs& a, hipStream_t st );
// Heat conduction over the bonds, then forward Euler of the temperature with dt.
void launchHeatTransfer( const PDArgs* args_dev, const PDArgs& a, double dt, hipStream_t st );
// dst[c * n + idx[m]] = a[m * ncomp + c] (+ b[m * ncomp + c] * t when b is set)
void launchPrescribed( const int64_t* idx_dev, int64_t count, const double* a_dev, const double* b_dev, int32_t ncomp,
                       double t, double* dst_dev, int64_t n, hipStream_t st );
// Packed launches over several specs of one kernel variant (the variant is
// taken from `variant`): tile t covers slices [tile_slice[t], +4) of
// specs[tile_spec[t]]; each spec's epilogue and displacement buffer follow
// from launch_index, its end_launch and parity_offset.
// The device launch counter of the batch engine's graph replays.
void launchAdvanceLaunch( int64_t* ctr, int64_t steps, hipStream_t st );
void launchSetLaunch( int64_t* ctr, int64_t L, hipStream_t st );
void launchForceBatched( const PDArgs* specs_dev, const PDArgs& variant, const int32_t* tile_spec,
                         const int64_t* tile_slice, int64_t ntiles, int64_t launch_index,
                         hipStream_t st );
void launchLpsStateBatched( const PDArgs* specs_dev, const PDArgs& variant, const int32_t* tile_spec,
                            const int64_t* tile_slice, int64_t ntiles, int64_t launch_index,
                            hipStream_t st );
// Whether the row-window force pass covers a specimen with these features
// (several materials as pair tables, a temperature, a plastic stretch).
bool rowFeaturesSupported( int32_t model, bool pair_tables, bool thermal, bool plastic );
const char* forceKernelName( const PDArgs& a );
const char* lpsKernelName( const PDArgs& a );
const char* energyKernelName( const PDArgs& a );

// ---- contact (pdk_contact.hip) -----------------------------------------------
void launchContactForce( const PDArgs* args_dev, const PDArgs& a, int32_t epilogue, int32_t parity,
                         hipStream_t st );
void launchContactInitialHalfStep( const PDArgs* args_dev, const PDArgs& a, int32_t parity,
                                   double* partial_dev, int64_t nblocks, hipStream_t st );
const char* contactKernelName( const PDArgs& a );

// ---- region outputs (pdk_output.hip) ----------------------------------------
enum RegionKind : int32_t
{
    kRegionAll = 0,
    kRegionBox = 1,
    kRegionCylinder = 2,
    kRegionLine = 3,
    kRegionIndices = 4
};
enum ReduceKind : int32_t
{
    kReduceSum = 0,
    kReduceMax = 1,
    kReduceMin = 2
};
struct RegionParams
{
    int32_t kind, dim;
    double low[3], high[3];
    double r_in, r_out, low_z, high_z, x_center, y_center;
    double half[2];
};
// Per-member value: src[p], zeroed where gate[p] <= threshold (gate set),
// times vol[p] (vol set).  src already points at the component's row.
struct QuantityArgs
{
    const double* src;
    const double* gate;
    const double* vol;
    double threshold;
};
constexpr int32_t kSeriesMaxPartials = 256;
void launchRegionMask( const RegionParams& R, const double* x_dev, const int64_t* inv_dev, int64_t n,
                       uint8_t* mask_dev, hipStream_t st );
int32_t seriesPartials( int64_t count );
void launchSeriesReduce( const int64_t* idx_dev, int64_t count, const QuantityArgs& Q, int32_t red,
                         double* partial_dev, double* out_dev, hipStream_t st );
void launchGatherQuantity( const int64_t* idx_dev, int64_t count, const QuantityArgs& Q, double* out_dev,
                           hipStream_t st );

} // namespace pdk

#endif
