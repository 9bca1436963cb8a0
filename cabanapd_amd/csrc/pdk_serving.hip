// pdk_serving.hip -- continuous batching of specimen solves on one GPU.
//
// Nothing in the reference corresponds to this: a CabanaPD run is one
// specimen per process.  Here every running specimen advances one time step
// per packed launch; all specimens of one kernel variant (model, fracture,
// precision, bond layout) share a launch whose tiles cover their concatenated slice
// ranges.  A specimen's epilogue (final half step only on its last step) and
// its displacement buffer follow from the launch index, so admission and
// retirement only rewrite the small tile tables, never the specimen state.
#include <algorithm>
#include <deque>
#include <limits>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "pdk_serving.hpp"

namespace pdk
{
struct BatchEngine::Impl
{
    struct Req
    {
        uint64_t ticket;
        pdk_spec_s* s;
        int32_t nsteps;
    };
    struct Run
    {
        uint64_t ticket;
        pdk_spec_s* s;
        int64_t end_launch;
        int32_t parity_offset;
        int32_t nsteps;
    };
    struct Group
    {
        PDArgs variant{};
        std::vector<PDArgs> h_specs;
        std::vector<int32_t> h_tspec;
        std::vector<int64_t> h_tslice;
        DevBuf specs, tspec, tslice;
        int64_t ntiles = 0;
        bool lps = false;
    };

    pdk_ctx_s* ctx;
    hipStream_t st = nullptr;
    int64_t L = 0;
    uint64_t next_ticket = 1;
    size_t max_slots = 512;
    // Admission keeps the running specimens within about this many slices
    // (enough wavefronts to fill the GPU several times over): arrivals are
    // admitted first come, first served, and small specimens share launches
    // while a large one runs nearly alone.
    int64_t max_slices = 16384;
    int64_t running_slices = 0;
    int32_t max_inflight = 16;  // packed steps queued ahead of the host
    std::deque<Req> pending;
    std::vector<Run> running;
    std::map<uint64_t, hipEvent_t> done_events;
    std::set<uint64_t> finished;
    std::deque<hipEvent_t> inflight;
    std::vector<Group> groups;
    bool dirty = true;
    // Graph replay of the packed step: kGraphSteps consecutive steps' launches
    // (gexec_n) and one step's (gexec_1) captured once per group
    // configuration, with the launch indexes counted from the device counter
    // ctr (each graph's last node advances it); replayed for up to
    // graph_steps steps at a time, never past the next retirement.  0 (the
    // default): eager launches only, which measure as fast here (the host
    // queues up to max_inflight steps ahead, so launch cost is hidden and a
    // packed step lasts as long as its kernels).
    static constexpr int64_t kGraphSteps = 8;
    int32_t graph_steps = 0;
    DevBuf ctr;
    hipGraphExec_t gexec_n = nullptr;
    hipGraphExec_t gexec_1 = nullptr;

    explicit Impl( pdk_ctx_s* c ) : ctx( c )
    {
        PDK_HIP( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
        int dev = 0, cus = 0;
        PDK_HIP( hipGetDevice( &dev ) );
        if ( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) == hipSuccess && cus > 0 )
            max_slices = static_cast<int64_t>( cus ) * 64;
    }
    ~Impl()
    {
        (void)hipStreamSynchronize( st );
        if ( gexec_n )
            (void)hipGraphExecDestroy( gexec_n );
        if ( gexec_1 )
            (void)hipGraphExecDestroy( gexec_1 );
        for ( auto& e : done_events )
            (void)hipEventDestroy( e.second );
        for ( auto e : inflight )
            (void)hipEventDestroy( e );
        (void)hipStreamDestroy( st );
    }

    void admit()
    {
        while ( !pending.empty() && running.size() < max_slots )
        {
            SpecView v = view( pending.front().s );
            if ( !running.empty() && running_slices + v.nslices > max_slices )
                break;
            Req r = pending.front();
            pending.pop_front();
            running_slices += v.nslices;
            // The spec's own stream may still be finishing its setup.
            hipEvent_t ready;
            PDK_HIP( hipEventCreateWithFlags( &ready, hipEventDisableTiming ) );
            PDK_HIP( hipEventRecord( ready, v.st ) );
            PDK_HIP( hipStreamWaitEvent( st, ready, 0 ) );
            PDK_HIP( hipEventDestroy( ready ) );
            // Initial half step of the first step (CabanaPD_Integrate.hpp:89-116).
            launchInitialHalfStep( v.dargs, *v.h, *v.cur, st );
            *v.cur ^= 1;
            Run run;
            run.ticket = r.ticket;
            run.s = r.s;
            run.nsteps = r.nsteps;
            run.end_launch = L + r.nsteps;
            run.parity_offset = *v.cur ^ static_cast<int32_t>( L & 1 );
            running.push_back( run );
            dirty = true;
        }
    }

    void rebuildGroups()
    {
        if ( !ctr )
            ctr.alloc( sizeof( int64_t ), &ctx_acct() ); // the device launch counter, on first use
        std::map<std::tuple<int, int, int, int>, size_t> key;
        std::vector<Group> ng;
        for ( const Run& r : running )
        {
            SpecView v = view( r.s );
            // one kernel variant per group: model, fracture, precision, bond layout
            const int form = ( useRowWindows( *v.h ) ? 2 : ( v.h->exist ? 1 : 0 ) ) +
                             ( v.h->rows != nullptr ? 4 : 0 ) + ( v.h->vol_uniform ? 0 : 8 ) +
                             ( v.h->any_nofail ? 16 : 0 );
            auto k = std::make_tuple( v.model, v.fracture, v.precision, form );
            auto it = key.find( k );
            size_t gi;
            if ( it == key.end() )
            {
                gi = ng.size();
                key[k] = gi;
                ng.emplace_back();
                ng[gi].variant = *v.h;
                ng[gi].lps = v.model == PDK_MODEL_LPS || v.model == PDK_MODEL_LINEAR_LPS;
            }
            else
                gi = it->second;
            Group& g = ng[gi];
            PDArgs a = *v.h;
            a.parity_offset = r.parity_offset;
            a.end_launch = r.end_launch;
            a.launch_ctr = ctr.as<int64_t>();
            const int32_t idx = static_cast<int32_t>( g.h_specs.size() );
            g.h_specs.push_back( a );
            for ( int64_t s0 = 0; s0 < v.nslices; s0 += kSlicesPerBlock )
            {
                g.h_tspec.push_back( idx );
                g.h_tslice.push_back( s0 );
            }
        }
        // Keep the old device tables alive until the stream has consumed them.
        PDK_HIP( hipStreamSynchronize( st ) );
        for ( Group& g : ng )
        {
            g.ntiles = static_cast<int64_t>( g.h_tspec.size() );
            g.specs.alloc( g.h_specs.size() * sizeof( PDArgs ), &ctx_acct() );
            g.tspec.alloc( g.h_tspec.size() * sizeof( int32_t ), &ctx_acct() );
            g.tslice.alloc( g.h_tslice.size() * sizeof( int64_t ), &ctx_acct() );
            PDK_HIP( hipMemcpyAsync( g.specs.as<PDArgs>(), g.h_specs.data(), g.h_specs.size() * sizeof( PDArgs ),
                                     hipMemcpyHostToDevice, st ) );
            PDK_HIP( hipMemcpyAsync( g.tspec.as<int32_t>(), g.h_tspec.data(), g.h_tspec.size() * sizeof( int32_t ),
                                     hipMemcpyHostToDevice, st ) );
            PDK_HIP( hipMemcpyAsync( g.tslice.as<int64_t>(), g.h_tslice.data(),
                                     g.h_tslice.size() * sizeof( int64_t ), hipMemcpyHostToDevice, st ) );
        }
        groups.swap( ng );
        dirty = false;
        dropGraph(); // the old graph's launches name the old tables
    }

    void dropGraph()
    {
        if ( gexec_n || gexec_1 )
            PDK_HIP( hipStreamSynchronize( st ) );
        for ( hipGraphExec_t* g : { &gexec_n, &gexec_1 } )
            if ( *g )
            {
                PDK_HIP( hipGraphExecDestroy( *g ) );
                *g = nullptr;
            }
    }

    // One packed step's launches in group order, at launch index L (eager) or
    // at the device counter's index + (-1 - L) (L < 0, a captured step).
    void launchStep( int64_t L )
    {
        for ( Group& g : groups )
        {
            if ( g.lps )
                launchLpsStateBatched( g.specs.as<PDArgs>(), g.variant, g.tspec.as<int32_t>(), g.tslice.as<int64_t>(),
                                       g.ntiles, L, st );
            launchForceBatched( g.specs.as<PDArgs>(), g.variant, g.tspec.as<int32_t>(), g.tslice.as<int64_t>(),
                                g.ntiles, L, st );
        }
    }

    // Capture `steps` consecutive packed steps and the counter's advance.
    hipGraphExec_t captureSteps( int64_t steps )
    {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        PDK_HIP( hipStreamBeginCapture( st, hipStreamCaptureModeThreadLocal ) );
        try
        {
            for ( int64_t j = 0; j < steps; j++ )
                launchStep( -1 - j );
            launchAdvanceLaunch( ctr.as<int64_t>(), steps, st );
        }
        catch ( ... )
        {
            (void)hipStreamEndCapture( st, &graph );
            if ( graph )
                (void)hipGraphDestroy( graph );
            throw;
        }
        PDK_HIP( hipStreamEndCapture( st, &graph ) );
        const hipError_t e = hipGraphInstantiate( &exec, graph, nullptr, nullptr, 0 );
        (void)hipGraphDestroy( graph );
        if ( e != hipSuccess )
            throw Error( std::string( "batch engine: graph instantiation failed: " ) + hipGetErrorString( e ) );
        return exec;
    }

    MemAccount& ctx_acct();

    void retire()
    {
        std::vector<Run> keep;
        hipEvent_t ev = nullptr;
        for ( const Run& r : running )
        {
            if ( r.end_launch != L )
            {
                keep.push_back( r );
                continue;
            }
            if ( !ev )
            {
                PDK_HIP( hipEventCreateWithFlags( &ev, hipEventDisableTiming ) );
                PDK_HIP( hipEventRecord( ev, st ) );
            }
            SpecView v = view( r.s );
            running_slices -= v.nslices;
            // The last launch (index end_launch - 1) did not flip the buffer.
            *v.cur = static_cast<int32_t>( ( ( r.end_launch - 1 ) & 1 ) ) ^ r.parity_offset;
            *v.steps_done += r.nsteps;
            *v.in_batch = false;
            PDK_HIP( hipStreamWaitEvent( v.st, ev, 0 ) );
            hipEvent_t mine;
            PDK_HIP( hipEventCreateWithFlags( &mine, hipEventDisableTiming ) );
            PDK_HIP( hipEventRecord( mine, st ) );
            done_events[r.ticket] = mine;
            dirty = true;
        }
        if ( ev )
            PDK_HIP( hipEventDestroy( ev ) );
        running.swap( keep );
    }

    int32_t progress( int32_t max_steps )
    {
        int32_t launched = 0;
        while ( max_steps <= 0 || launched < max_steps )
        {
            admit();
            if ( running.empty() )
                break;
            if ( dirty )
                rebuildGroups();
            // steps until the next retirement (arrivals wait for one: admit()
            // has taken every pending spec that fits)
            int64_t k = std::numeric_limits<int64_t>::max();
            for ( const Run& r : running )
                k = std::min( k, r.end_launch - L );
            if ( max_steps > 0 )
                k = std::min<int64_t>( k, max_steps - launched );
            k = std::min<int64_t>( k, graph_steps );
            if ( k >= 2 )
            {
                if ( k >= kGraphSteps && !gexec_n )
                    gexec_n = captureSteps( kGraphSteps );
                if ( k % kGraphSteps != 0 && !gexec_1 )
                    gexec_1 = captureSteps( 1 );
                launchSetLaunch( ctr.as<int64_t>(), L, st );
                for ( int64_t i = 0; i < k / kGraphSteps; i++ )
                    PDK_HIP( hipGraphLaunch( gexec_n, st ) );
                for ( int64_t i = 0; i < k % kGraphSteps; i++ )
                    PDK_HIP( hipGraphLaunch( gexec_1, st ) );
            }
            else
            {
                k = 1;
                launchStep( L );
            }
            L += k;
            retire();
            launched += static_cast<int32_t>( k );
            hipEvent_t e;
            PDK_HIP( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
            PDK_HIP( hipEventRecord( e, st ) );
            inflight.push_back( e );
            while ( inflight.size() > static_cast<size_t>( max_inflight ) )
            {
                PDK_HIP( hipEventSynchronize( inflight.front() ) );
                PDK_HIP( hipEventDestroy( inflight.front() ) );
                inflight.pop_front();
            }
        }
        return launched;
    }

    bool done( uint64_t t )
    {
        if ( finished.count( t ) )
            return true;
        auto it = done_events.find( t );
        if ( it == done_events.end() )
            return false;
        hipError_t q = hipEventQuery( it->second );
        if ( q == hipSuccess )
        {
            (void)hipEventDestroy( it->second );
            done_events.erase( it );
            finished.insert( t );
            return true;
        }
        if ( q != hipErrorNotReady )
            throw Error( std::string( "batch engine: " ) + hipGetErrorString( q ) );
        return false;
    }

    void forget( pdk_spec_s* s )
    {
        PDK_HIP( hipStreamSynchronize( st ) );
        pending.erase( std::remove_if( pending.begin(), pending.end(), [&]( const Req& r ) { return r.s == s; } ),
                       pending.end() );
        size_t before = running.size();
        for ( const Run& r : running )
            if ( r.s == s )
                running_slices -= view( s ).nslices;
        running.erase( std::remove_if( running.begin(), running.end(), [&]( const Run& r ) { return r.s == s; } ),
                       running.end() );
        if ( running.size() != before )
            dirty = true;
    }
};

} // namespace pdk

// The context owns the accounting record; reach it through the ctx pointer.
namespace pdk
{
MemAccount& ctxAccount( pdk_ctx_s* c );
MemAccount& BatchEngine::Impl::ctx_acct()
{
    return ctxAccount( ctx );
}

BatchEngine::BatchEngine( pdk_ctx_s* ctx ) : impl_( std::make_unique<Impl>( ctx ) ) {}
BatchEngine::~BatchEngine() = default;

uint64_t BatchEngine::submit( pdk_spec_s* s, int32_t nsteps )
{
    SpecView v = view( s );
    if ( !v.inited )
        throw Error( "call pdk_init before pdk_submit" );
    if ( *v.in_batch )
        throw Error( "spec already submitted" );
    if ( v.model == PDK_MODEL_HERTZ || v.model == PDK_MODEL_HERTZ_JKR )
        throw Error( "the batch engine runs PD specimens only" );
    if ( nsteps <= 0 )
        throw Error( "nsteps must be positive" );
    *v.in_batch = true;
    const uint64_t t = impl_->next_ticket++;
    impl_->pending.push_back( { t, s, nsteps } );
    return t;
}

bool BatchEngine::done( uint64_t ticket ) const
{
    return impl_->done( ticket );
}

int32_t BatchEngine::progress( int32_t max_steps )
{
    return impl_->progress( max_steps );
}

void BatchEngine::forget( pdk_spec_s* s )
{
    impl_->forget( s );
}

void BatchEngine::setGraphSteps( int32_t steps )
{
    if ( steps < 0 )
        throw Error( "graph steps must be 0 (eager) or positive" );
    impl_->dropGraph();
    impl_->graph_steps = steps;
}

} // namespace pdk
