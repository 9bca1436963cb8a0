// pdk_serving.hpp -- internal interfaces between the C ABI translation unit,
// the contact step and the continuous-batching engine.
#ifndef PDK_SERVING_HPP
#define PDK_SERVING_HPP

#include <functional>
#include <memory>

#include "pdk_internal.hpp"

struct pdk_ctx_s;
struct pdk_spec_s;

namespace pdk
{
// Read-mostly view of a spec for the batch engine.
struct SpecView
{
    PDArgs* h;
    const PDArgs* dargs;
    hipStream_t st;
    int64_t n, nslices;
    int32_t* cur;
    int64_t* steps_done;
    bool* in_batch;
    bool inited;
    int32_t model, fracture, precision;
    uint64_t bytes;
};
SpecView view( pdk_spec_s* s );
void specPushArgs( pdk_spec_s* s );
void specTimed( pdk_spec_s* s, int cls, const std::function<void()>& fn );

struct ContactState
{
    double* max_disp;
    int64_t* since_build;
    double extend;
    int32_t rebuild_every;
    DevBuf* partial;
    MemAccount* acct;
    double* disp_pinned;   // pinned host copy of the largest displacement
    hipEvent_t disp_ready; // recorded after that copy
};
ContactState contactState( pdk_spec_s* s );
void contactBuildImpl( pdk_spec_s* s, bool initial, double* max_disp, int64_t* since_build );

// Continuous batching over one device: specs submitted with a step count are
// admitted while slots and the HBM budget allow, and every packed step
// advances all running specs with one launch per kernel variant over their
// concatenated slice ranges.
class BatchEngine
{
  public:
    explicit BatchEngine( pdk_ctx_s* ctx );
    ~BatchEngine();
    uint64_t submit( pdk_spec_s* s, int32_t nsteps );
    bool done( uint64_t ticket ) const;
    int32_t progress( int32_t max_steps );
    void forget( pdk_spec_s* s );
    // Replay the captured packed step for up to steps steps at a time (0: eager).
    void setGraphSteps( int32_t steps );

  private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

} // namespace pdk

#endif
