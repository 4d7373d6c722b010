// Render server: camera sessions over shared static scenes, with the frames
// of many sessions in flight together on a few HIP streams.
//
// A session is the per-camera state of the reference's VulkanDevice (the
// A-SVGF history and moments, ReSTIR DI/GI reservoirs, the light grid and the
// adapted luminance), kept resident in HBM; a frame is the part of
// VulkanDevice::Render (Source/VulkanDevice.cpp:540-588) from the traced
// passes to the final compose, as EncodeFrames runs it.  Within-frame planes
// and wavefront buffers belong to a stream, so a session costs only what it
// carries from frame to frame.  One launch renders a batch of waiting
// requests of sessions of one scene, or of the scenes of one scene group:
// every traversal stage of the batch is one persistent launch over all its
// frames' rays (a group's rays each start at their own scene's root of the
// group's batched BVH).
#include "rg_frame.h"
#include "rg_internal.h"
#include "rg_kernels.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <optional>
#include <thread>
#include <vector>

namespace rg
{
namespace
{

double NowSeconds()
{
    return std::chrono::duration< double >( std::chrono::steady_clock::now().time_since_epoch() ).count();
}

// The parameter structures a frame reads, copied at submit so that the
// caller's chain may go away before the frame runs.
struct FrameRequest
{
    RgDrawFrameInfo               di{};
    RgDrawFrameTonemappingParams  tm{};
    RgDrawFrameSkyParams          sky{};
    RgDrawFrameIlluminationParams il{};
    RgDrawFrameTexturesParams     tx{};
    RgExtDrawFramePathParams      path{};
    RgDrawFrameReflectRefractParams rr{};
    bool                          hasTm = false, hasSky = false, hasIl = false, hasTx = false, hasPath = false, hasRr = false;
    uint64_t                      tag    = 0;
    double                        submit = 0.0;

    // What frames of one batch must share: the ray cull masks and the
    // reflection / refraction depth (FillFrameParams derives the rest of
    // TraceCompatible's fields from the scene).
    uint64_t BatchKey() const
    {
        const uint32_t depth = std::min( 16u, hasRr ? rr.maxReflectRefractDepth : 2u );
        return ( uint64_t( di.rayCullMaskWorld ) << 32 ) | depth;
    }

    void Capture( const RgDrawFrameInfo& in )
    {
        di = in;
        if( const auto* p = FindParams< RgDrawFrameTonemappingParams >( in, RG_STRUCTURE_TYPE_TONEMAPPING ) )
        {
            tm = *p, hasTm = true;
        }
        if( const auto* p = FindParams< RgDrawFrameSkyParams >( in, RG_STRUCTURE_TYPE_SKY ) )
        {
            sky = *p, hasSky = true;
            sky.pSkyCubemapTextureName = nullptr;
        }
        if( const auto* p = FindParams< RgDrawFrameIlluminationParams >( in, RG_STRUCTURE_TYPE_ILLUMINATION ) )
        {
            il = *p, hasIl = true;
            // the scene's lights are encoded once, so light styles do not apply
            il.lightUniqueIdIgnoreFirstPersonViewerShadows = nullptr;
            il.lightstyleValuesCount                       = 0;
            il.pLightstyleValues                           = nullptr;
        }
        if( const auto* p = FindParams< RgDrawFrameTexturesParams >( in, RG_STRUCTURE_TYPE_TEXTURES ) )
        {
            tx = *p, hasTx = true;
        }
        if( const auto* p = FindParams< RgExtDrawFramePathParams >( in, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS ) )
        {
            path = *p, hasPath = true;
        }
        if( const auto* p = FindParams< RgDrawFrameReflectRefractParams >( in, RG_STRUCTURE_TYPE_REFLECTREFRACT ) )
        {
            rr = *p, hasRr = true;
        }
        Relink();
    }
    // Rebuilds di.pParams over the copies (after the request has moved).
    void Relink()
    {
        void* head = nullptr;
        auto  push = [ & ]( bool has, auto& s ) {
            if( has )
            {
                s.pNext = head;
                head    = &s;
            }
        };
        push( hasTm, tm );
        push( hasSky, sky );
        push( hasIl, il );
        push( hasTx, tx );
        push( hasPath, path );
        push( hasRr, rr );
        di.pParams = head;
    }
};

} // namespace
} // namespace rg

using namespace rg;

struct RgExtServer_T;
struct RgExtSceneGroup_T;

struct RgExtScene_T
{
    RgExtServer_T*             server = nullptr;
    SceneCollector          sc;
    DeviceScene             ds;
    BvhObject               bvh;
    std::vector< uint32_t > geomMap;
    bool                    built    = false;
    uint32_t                sessions = 0;
    size_t                  bytes    = 0;
    RgExtSceneGroup_T*      group      = nullptr; // traced through the group's BVH while set
    uint32_t                groupIndex = 0;       // the scene's number in the group's BVH
    bool                    joining    = false;   // a group with it is being built
};

// Scenes whose frames share launches: one batched BVH over their triangles
// (scene k = subtree k).
struct RgExtSceneGroup_T
{
    RgExtServer_T*                server = nullptr;
    std::vector< RgExtScene_T* > members;
    BvhObject                    bvh;
    size_t                       bytes = 0;
};

struct RgExtSession_T
{
    RgExtServer_T*                   server = nullptr;
    RgExtScene_T*                    scene  = nullptr;
    uint32_t                      width = 0, height = 0, stream = 0;
    RgExtFramebuffers_T           fbs;
    FrameGpu                      g;
    CameraState                   cam;
    std::optional< FrameRequest > pending;
    bool                          queued   = false;
    bool                          dying    = false;
    uint32_t                      inflight = 0;
    size_t                        bytes    = 0;
    void*                         image[ 2 ] = { nullptr, nullptr }; // pinned RGBA8 readback images, used in turn
    uint32_t                      imageNext  = 0;

    ~RgExtSession_T()
    {
        for( void* p : image )
        {
            if( p )
            {
                ( void )hipHostFree( p );
            }
        }
    }
};

namespace rg
{
namespace
{
struct StreamCtx
{
    struct Member
    {
        RgExtSession_T* s;
        uint64_t        tag;
        uint32_t        frameId;
        double          submit;
        const void*     image;
    };
    // One launch: the frames of a batch, finished together.
    struct InFlight
    {
        std::vector< Member > m;
        double                start;
        hipEvent_t            ev;
    };
    hipStream_t                    stream = nullptr;
    std::vector< std::unique_ptr< TransientPlanes > > planes; // one set per frame of a batch
    Wavefront                      wf;     // every frame of a batch at the largest render size
    TraceContext                   tc;
    DeviceBuffer                   packed; // RGBA8 of a batch's frames for the readback
    std::vector< hipEvent_t > events;
    uint32_t                  nextEvent = 0;
    std::deque< InFlight >    inflight;

    ~StreamCtx()
    {
        for( hipEvent_t e : events )
        {
            ( void )hipEventDestroy( e );
        }
        if( stream )
        {
            ( void )hipStreamDestroy( stream );
        }
    }
};
} // namespace
} // namespace rg

struct RgExtServer_T
{
    RgExtServerCreateInfo                       info{};
    uint32_t                                    nodeFormat = RG_EXT_BVH_NODE_FP32;
    uint32_t                                    batchCapacity = 1; // within-frame plane sets per stream
    uint32_t                                    activeStreams = 1; // streams that take an idle session's next launch
    uint32_t                                    storage    = RG_EXT_STORAGE_FP32;
    hipStream_t                                 setup      = nullptr; // scene builds and session set-up
    std::vector< std::unique_ptr< StreamCtx > > streams;
    std::list< std::unique_ptr< RgExtScene_T > >   scenes;
    std::list< std::unique_ptr< RgExtSceneGroup_T > > groups;
    std::list< std::unique_ptr< RgExtSession_T > > sessions;
    std::deque< RgExtSession_T* >                  ready; // sessions with a waiting request
    std::deque< RgExtFrameResult >              results;
    std::mutex                                  m;
    std::condition_variable                     cvWork, cvIdle;
    std::thread                                 worker;
    bool                                        stop = false;
    uint32_t                                    launching  = 0; // frames being encoded outside the lock
    uint32_t                                    roundRobin = 0;
    uint64_t                                    sessionBytes = 0, sharedBytes = 0, sceneBytes = 0;
    uint64_t                                    done = 0, dropped = 0, failed = 0;

    void Result( RgExtSession_T* s, uint64_t tag, uint32_t status, uint32_t frameId, double submit, double start, double end,
                 const void* image = nullptr, uint32_t batch = 0 )
    {
        RgExtFrameResult r{};
        r.session    = reinterpret_cast< RgExtSession >( s );
        r.tag        = tag;
        r.status     = status;
        r.frameId    = frameId;
        r.submitTime = submit;
        r.startTime  = start;
        r.endTime    = end;
        r.pImage     = image;
        r.batchSize  = batch;
        results.push_back( r );
        ( status == RG_EXT_FRAME_DONE ? done : status == RG_EXT_FRAME_FAILED ? failed : dropped )++;
    }

    bool Busy() const
    {
        if( !ready.empty() || launching > 0 )
        {
            return true;
        }
        for( const auto& st : streams )
        {
            if( !st->inflight.empty() )
            {
                return true;
            }
        }
        return false;
    }

    // Retires the frames whose completion events have fired.
    void Retire()
    {
        for( auto& st : streams )
        {
            while( !st->inflight.empty() )
            {
                const auto&      f = st->inflight.front();
                const hipError_t q = hipEventQuery( f.ev );
                if( q == hipErrorNotReady )
                {
                    break;
                }
                const double end = NowSeconds();
                for( const auto& m : f.m )
                {
                    Result( m.s, m.tag, q == hipSuccess ? RG_EXT_FRAME_DONE : RG_EXT_FRAME_FAILED, m.frameId, m.submit, f.start, end, m.image,
                            uint32_t( f.m.size() ) );
                    m.s->inflight--;
                }
                st->inflight.pop_front();
            }
        }
    }

    // Drops waiting requests older than maxQueueSeconds.
    void Expire()
    {
        if( info.maxQueueSeconds <= 0.0 )
        {
            return;
        }
        const double now = NowSeconds();
        for( auto it = ready.begin(); it != ready.end(); )
        {
            RgExtSession_T* s = *it;
            if( now - s->pending->submit > info.maxQueueSeconds )
            {
                Result( s, s->pending->tag, RG_EXT_FRAME_EXPIRED, 0, s->pending->submit, 0.0, now );
                s->pending.reset();
                s->queued = false;
                it        = ready.erase( it );
            }
            else
            {
                ++it;
            }
        }
    }

    // The stream a session's next frame goes to: its own while a frame of
    // it is in flight (its frames must stay ordered), else the one with the
    // fewest launches queued (its persistent planes are its own, the
    // within-frame planes are taken from that stream).
    uint32_t TargetStream( const RgExtSession_T& s ) const
    {
        uint32_t target = s.stream;
        if( s.inflight == 0 )
        {
            if( target >= activeStreams )
            {
                target = 0;
            }
            for( uint32_t k = 0; k < activeStreams; k++ )
            {
                if( streams[ k ]->inflight.size() < streams[ target ]->inflight.size() )
                {
                    target = k;
                }
            }
        }
        return target;
    }

    // Called with the lock held: picks the next batch (the lead request by
    // the schedule, then up to batchFrames - 1 more waiting requests of the
    // same scene, or of a scene of the lead scene's group, and batch key
    // that may run on the lead's stream, oldest first) and launches it.
    // False when no stream has room for a lead.
    bool LaunchNext( std::unique_lock< std::mutex >& lk )
    {
        auto lead = ready.end();
        for( auto it = ready.begin(); it != ready.end(); ++it )
        {
            if( streams[ TargetStream( **it ) ]->inflight.size() >= info.framesInFlightPerStream )
            {
                continue;
            }
            if( lead == ready.end() ||
                ( info.schedule == RG_EXT_SCHEDULE_NEWEST_FIRST ? ( *it )->pending->submit > ( *lead )->pending->submit
                                                                 : ( *it )->pending->submit < ( *lead )->pending->submit ) )
            {
                lead = it;
            }
        }
        if( lead == ready.end() )
        {
            return false;
        }
        RgExtSession_T*                s0     = *lead;
        const uint32_t                 target = TargetStream( *s0 );
        const uint64_t                 key    = s0->pending->BatchKey();
        // a frame that culls world 0 sees nothing of a grouped scene: it
        // traces the scene's own BVH, which drops that class the usual way
        RgExtSceneGroup_T* const grp =
            ( s0->pending->di.rayCullMaskWorld & RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT ) ? s0->scene->group : nullptr;
        std::vector< RgExtSession_T* > batch{ s0 };
        ready.erase( lead );
        if( info.batchFrames > 1 )
        {
            std::vector< RgExtSession_T* > cand;
            for( RgExtSession_T* s : ready )
            {
                const bool sameTrace = grp ? s->scene->group == grp : s->scene == s0->scene;
                if( sameTrace && s->pending->BatchKey() == key && ( s->inflight == 0 || s->stream == target ) )
                {
                    cand.push_back( s );
                }
            }
            std::stable_sort( cand.begin(), cand.end(),
                              []( const RgExtSession_T* a, const RgExtSession_T* b ) { return a->pending->submit < b->pending->submit; } );
            for( RgExtSession_T* s : cand )
            {
                if( batch.size() >= info.batchFrames )
                {
                    break;
                }
                batch.push_back( s );
                ready.erase( std::find( ready.begin(), ready.end(), s ) );
            }
        }
        if( grp )
        {
            // a grouped launch takes each scene's rays as one range
            std::stable_sort( batch.begin(), batch.end(),
                              []( const RgExtSession_T* a, const RgExtSession_T* b ) { return a->scene->groupIndex < b->scene->groupIndex; } );
        }
        LaunchBatch( batch, grp, *streams[ target ], target, lk );
        return true;
    }

    // Called with the lock held; encodes the batch with the lock released,
    // so submissions, polls and statistics do not wait for the host-side
    // encoding.  Each session counts its frame as in flight meanwhile (a
    // destroy waits for it) and the server stays busy (WaitIdle waits).
    // Only this thread launches frames and touches the sessions' cameras,
    // framebuffers and the stream's state.  grp: the group whose BVH the
    // batch traces (its scenes' frames, in group order), or null for the
    // frames of one ungrouped scene; a group outlives the launches that
    // took it (rgExtSceneGroupDestroy waits for them).
    void LaunchBatch( const std::vector< RgExtSession_T* >& batch, RgExtSceneGroup_T* grp, StreamCtx& st, uint32_t target,
                      std::unique_lock< std::mutex >& lk )
    {
        const uint32_t              B = uint32_t( batch.size() );
        std::vector< FrameRequest > req( B );
        for( uint32_t k = 0; k < B; k++ )
        {
            RgExtSession_T& s = *batch[ k ];
            req[ k ]          = std::move( *s.pending );
            s.pending.reset();
            s.queued = false;
            req[ k ].Relink();
            s.inflight++;
            s.stream = target;
        }
        launching++;
        lk.unlock();
        const double                     start = NowSeconds();
        hipEvent_t                       ev    = nullptr;
        bool                             ok    = false;
        std::string                      failure;
        std::vector< StreamCtx::Member > members;
        try
        {
            EncodeSlot slots[ MAX_TRACE_SLOTS ];
            for( uint32_t k = 0; k < B; k++ )
            {
                RgExtSession_T& s  = *batch[ k ];
                RgExtScene_T&   sc = *s.scene;
                FillFrameParams( s.cam, req[ k ].di, s.width, s.height, uint32_t( sc.sc.lights.size() ), uint32_t( sc.sc.lights.size() ),
                                 sc.sc.hasDirectional, false );
                if( !s.cam.firstFrame )
                {
                    s.fbs.fb.Swap();
                }
                s.fbs.fb.shared = st.planes[ k ].get();
                slots[ k ]      = EncodeSlot{ &s.g, &s.cam.params, &s.fbs.fb, &req[ k ].di, &sc.ds, sc.groupIndex };
            }
            if( grp )
            {
                EncodeFrames( slots, B, st.tc, grp->bvh, nullptr, st.wf, st.stream );
            }
            else
            {
                EncodeFrames( slots, B, st.tc, batch[ 0 ]->scene->bvh, &batch[ 0 ]->scene->ds, st.wf, st.stream );
            }
            size_t packOff = 0;
            for( uint32_t k = 0; k < B; k++ )
            {
                RgExtSession_T& s     = *batch[ k ];
                const void*     image = nullptr;
                if( info.flags & RG_EXT_SERVER_READBACK )
                {
                    // FINAL -> RGBA8 in the stream's staging, then to the
                    // session's pinned image of this frame
                    const size_t n = size_t( s.width ) * s.height;
                    uint32_t*    d = st.packed.as< uint32_t >() + packOff;
                    PackFinalRgba8( reinterpret_cast< const float* >( s.fbs.fb.Cur( RG_EXT_PLANE_FINAL ) ), d, uint32_t( n ), st.stream );
                    void* h = s.image[ s.imageNext ];
                    s.imageNext ^= 1u;
                    RG_HIP_CHECK( hipMemcpyAsync( h, d, n * 4, hipMemcpyDeviceToHost, st.stream ) );
                    image = h;
                    packOff += n;
                }
                members.push_back( { &s, req[ k ].tag, s.cam.frameId, req[ k ].submit, image } );
            }
            ev = st.events[ st.nextEvent++ % st.events.size() ];
            RG_HIP_CHECK( hipEventRecord( ev, st.stream ) );
            ok = true;
        }
        catch( const Error& e )
        {
            failure = e.what();
        }
        lk.lock();
        launching--;
        if( ok )
        {
            st.inflight.push_back( { std::move( members ), start, ev } );
            for( RgExtSession_T* s : batch )
            {
                s->cam.frameId++;
                s->cam.firstFrame = false;
            }
        }
        else
        {
            SetLastError( failure.c_str() );
            for( uint32_t k = 0; k < B; k++ )
            {
                batch[ k ]->inflight--;
                Result( batch[ k ], req[ k ].tag, RG_EXT_FRAME_FAILED, batch[ k ]->cam.frameId, req[ k ].submit, start, NowSeconds() );
            }
        }
    }

    void Run()
    {
        std::unique_lock< std::mutex > lk( m );
        for( ;; )
        {
            Retire();
            Expire();
            bool launched = false;
            while( !ready.empty() && LaunchNext( lk ) )
            {
                launched = true;
                Retire();
            }
            cvIdle.notify_all();
            if( stop && !Busy() )
            {
                return;
            }
            if( !launched )
            {
                bool inflight = false;
                for( const auto& st : streams )
                {
                    inflight = inflight || !st->inflight.empty();
                }
                if( inflight )
                {
                    cvWork.wait_for( lk, std::chrono::microseconds( 50 ) );
                }
                else if( info.maxQueueSeconds > 0.0 && !ready.empty() )
                {
                    cvWork.wait_for( lk, std::chrono::milliseconds( 1 ) );
                }
                else
                {
                    cvWork.wait( lk, [ & ] { return stop || !ready.empty(); } );
                }
            }
        }
    }

    void WaitIdle()
    {
        std::unique_lock< std::mutex > lk( m );
        cvIdle.wait( lk, [ & ] { return !Busy(); } );
    }

    ~RgExtServer_T()
    {
        if( worker.joinable() )
        {
            {
                std::lock_guard< std::mutex > lk( m );
                stop = true;
            }
            cvWork.notify_all();
            worker.join();
        }
        sessions.clear();
        groups.clear();
        scenes.clear();
        streams.clear();
        if( setup )
        {
            ( void )hipStreamDestroy( setup );
        }
    }
};

namespace
{
RgExtServer_T& Srv( RgExtServer s )
{
    RG_REQUIRE( s, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "server is null" );
    return *s;
}
} // namespace

extern "C" {

RgResult rgExtServerCreate( const RgExtServerCreateInfo* pInfo, RgExtServer* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pInfo && pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RgExtServerCreateInfo ci = *pInfo;
        ci.streamCount             = ci.streamCount ? ci.streamCount : 4u;
        ci.framesInFlightPerStream = ci.framesInFlightPerStream ? ci.framesInFlightPerStream : 2u;
        ci.batchFrames             = ci.batchFrames ? ci.batchFrames : 1u;
        RG_REQUIRE( ci.streamCount <= 16 && ci.framesInFlightPerStream <= 8, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "streamCount must be <= 16 and framesInFlightPerStream <= 8" );
        RG_REQUIRE( ci.batchFrames <= MAX_TRACE_SLOTS, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "batchFrames must be <= 16" );
        RG_REQUIRE( ci.schedule <= RG_EXT_SCHEDULE_NEWEST_FIRST, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown schedule" );
        RG_REQUIRE( ( ci.flags & ~RG_EXT_SERVER_READBACK ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown server flags" );
        RG_REQUIRE( ci.maxQueueSeconds >= 0.0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "maxQueueSeconds must not be negative" );
        RG_REQUIRE( ( ci.precisionFlags & ~( RG_EXT_PRECISION_FP16_STORAGE | RG_EXT_PRECISION_FP16_NODES ) ) == 0,
                    RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown precisionFlags bits" );
        auto srv        = std::make_unique< RgExtServer_T >();
        srv->info       = ci;
        srv->batchCapacity = ci.batchFrames;
        srv->activeStreams = ci.streamCount;
        srv->storage    = ( ci.precisionFlags & RG_EXT_PRECISION_FP16_STORAGE ) ? RG_EXT_STORAGE_FP16 : RG_EXT_STORAGE_FP32;
        srv->nodeFormat = ( ci.precisionFlags & RG_EXT_PRECISION_FP16_NODES ) ? RG_EXT_BVH_NODE_FP16 : RG_EXT_BVH_NODE_FP32;
        RG_HIP_CHECK( hipStreamCreateWithFlags( &srv->setup, hipStreamNonBlocking ) );
        const size_t capPx = size_t( ci.maxRenderWidth ) * ci.maxRenderHeight;
        for( uint32_t i = 0; i < ci.streamCount; i++ )
        {
            auto st = std::make_unique< StreamCtx >();
            RG_HIP_CHECK( hipStreamCreateWithFlags( &st->stream, hipStreamNonBlocking ) );
            for( uint32_t k = 0; k < ci.batchFrames; k++ )
            {
                st->planes.push_back( std::make_unique< TransientPlanes >() );
                st->planes.back()->Create( ci.maxRenderWidth, ci.maxRenderHeight, srv->storage );
            }
            st->wf.Reserve( capPx * ci.batchFrames );
            if( ci.flags & RG_EXT_SERVER_READBACK )
            {
                st->packed.Reserve( capPx * ci.batchFrames * 4 );
            }
            st->events.resize( ci.framesInFlightPerStream );
            for( auto& e : st->events )
            {
                RG_HIP_CHECK( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
            }
            srv->sharedBytes += ( FramebufferSet::BytesFor( ci.maxRenderWidth, ci.maxRenderHeight, srv->storage, false ) -
                                  FramebufferSet::BytesFor( ci.maxRenderWidth, ci.maxRenderHeight, srv->storage, true ) + Wavefront::BytesFor( capPx ) ) *
                                    ci.batchFrames +
                                st->packed.bytes;
            srv->streams.push_back( std::move( st ) );
        }
        RgExtServer_T* raw = srv.get();
        srv->worker     = std::thread( [ raw ] { raw->Run(); } );
        *pOut           = srv.release();
    } );
}

RgResult rgExtServerDestroy( RgExtServer server )
{
    return Guard( [ & ] {
        Srv( server ).WaitIdle();
        delete server;
    } );
}

RgResult rgExtSceneCreate( RgExtServer server, RgExtScene* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        RgExtServer_T& s  = Srv( server );
        auto        sc = std::make_unique< RgExtScene_T >();
        sc->server     = &s;
        sc->sc.BeginFrame();
        RgExtScene out = sc.get();
        std::lock_guard< std::mutex > lk( s.m );
        s.scenes.push_back( std::move( sc ) );
        *pOut = out;
    } );
}

RgResult rgExtSceneUploadMeshPrimitive( RgExtScene scene, const RgMeshInfo* pMesh, const RgMeshPrimitiveInfo* pPrimitive )
{
    return Guard( [ & ] {
        RG_REQUIRE( scene && pMesh && pPrimitive, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( !scene->built, RG_RESULT_WRONG_FUNCTION_CALL, "the scene is already built" );
        scene->sc.AddPrimitive( *pMesh, *pPrimitive );
    } );
}

RgResult rgExtSceneUploadDirectionalLight( RgExtScene scene, const RgDirectionalLightUploadInfo* pInfo )
{
    return Guard( [ & ] {
        RG_REQUIRE( scene && pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( !scene->built, RG_RESULT_WRONG_FUNCTION_CALL, "the scene is already built" );
        scene->sc.AddDirectional( *pInfo );
    } );
}

RgResult rgExtSceneUploadSphericalLight( RgExtScene scene, const RgSphericalLightUploadInfo* pInfo )
{
    return Guard( [ & ] {
        RG_REQUIRE( scene && pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( !scene->built, RG_RESULT_WRONG_FUNCTION_CALL, "the scene is already built" );
        scene->sc.AddSphere( *pInfo );
    } );
}

RgResult rgExtSceneBuild( RgExtScene scene )
{
    return Guard( [ & ] {
        RG_REQUIRE( scene, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "scene is null" );
        RG_REQUIRE( !scene->built, RG_RESULT_WRONG_FUNCTION_CALL, "the scene is already built" );
        RgExtServer_T&     s  = *scene->server;
        SceneCollector& sc = scene->sc;
        // A static scene is its own previous frame: every light and geometry
        // maps to itself (the steady state of an instance that re-uploads the
        // same scene each frame).
        sc.prevLightIndex.clear();
        for( uint32_t i = 0; i < sc.lights.size(); i++ )
        {
            if( sc.lights[ i ].lightType != LIGHT_TYPE_NONE )
            {
                sc.prevLightIndex[ sc.lightIds[ i ] ] = i;
            }
        }
        sc.prevLightCount = uint32_t( sc.lights.size() );
        sc.FinishLights();
        sc.Seal();
        sc.prevGeomUids = sc.geomUids;
        sc.GeomPrevToCur( scene->geomMap );
        UploadScene( sc, scene->ds, s.setup );
        AssembleScene( scene->ds, s.setup );
        if( scene->ds.triCount > 0 )
        {
            RgExtBvhBuildParams bp = { 1, s.nodeFormat };
            BuildFrameBvh( scene->bvh, scene->ds.tris.as< float >(), scene->ds.classTris, GC_COUNT, bp, s.setup );
            scene->ds.bvhValid      = true;
            scene->ds.builtTriCount = scene->ds.triCount;
        }
        RG_HIP_CHECK( hipStreamSynchronize( s.setup ) );
        sc.vertices.Free();
        sc.indices.Free();
        const size_t n = scene->ds.triCount;
        scene->bytes   = n * ( 9 + 9 + 1 + 9 ) * 4 + size_t( scene->bvh.nodeCount ) * scene->bvh.NodeStride() + n * 48;
        scene->built   = true;
        std::lock_guard< std::mutex > lk( s.m );
        s.sceneBytes += scene->bytes;
    } );
}

RgResult rgExtSceneDestroy( RgExtScene scene )
{
    return Guard( [ & ] {
        RG_REQUIRE( scene, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "scene is null" );
        RgExtServer_T&                   s = *scene->server;
        std::unique_ptr< RgExtScene_T >  keep;
        {
            std::lock_guard< std::mutex > lk( s.m );
            RG_REQUIRE( scene->sessions == 0, RG_RESULT_WRONG_FUNCTION_CALL, "sessions still use the scene" );
            RG_REQUIRE( !scene->group && !scene->joining, RG_RESULT_WRONG_FUNCTION_CALL, "the scene is a member of a group" );
            for( auto it = s.scenes.begin(); it != s.scenes.end(); ++it )
            {
                if( it->get() == scene )
                {
                    keep = std::move( *it );
                    s.scenes.erase( it );
                    break;
                }
            }
            s.sceneBytes -= scene->bytes;
        }
    } );
}

RgResult rgExtSceneGroupCreate( RgExtServer server, const RgExtScene* pScenes, uint32_t sceneCount, RgExtSceneGroup* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pScenes && pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RgExtServer_T& s = Srv( server );
        RG_REQUIRE( sceneCount >= 2 && sceneCount <= 128, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a group has 2..128 scenes" );
        auto grp    = std::make_unique< RgExtSceneGroup_T >();
        grp->server = &s;
        std::vector< uint32_t > offsets( 1, 0u );
        {
            std::lock_guard< std::mutex > lk( s.m );
            for( uint32_t k = 0; k < sceneCount; k++ )
            {
                RgExtScene_T* sc = pScenes[ k ];
                RG_REQUIRE( sc && sc->server == &s, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a scene is null or of another server" );
                RG_REQUIRE( sc->built, RG_RESULT_WRONG_FUNCTION_CALL, "a scene is not built" );
                RG_REQUIRE( !sc->group && !sc->joining, RG_RESULT_WRONG_FUNCTION_CALL, "a scene is already in a group" );
                RG_REQUIRE( std::find( grp->members.begin(), grp->members.end(), sc ) == grp->members.end(), RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                            "a scene is listed twice" );
                RG_REQUIRE( GroupableScene( sc->ds ), RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                            "a scene has primitives other than opaque world-0 geometry" );
                RG_REQUIRE( sc->ds.triCount >= 2, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a scene has fewer than 2 triangles" );
                const RgExtScene_T& first = *pScenes[ 0 ];
                // lights[0] is the directional light's slot, the rest are
                // the other lights (TraceCompatible's light kinds)
                RG_REQUIRE( sc->sc.hasDirectional == first.sc.hasDirectional && ( sc->sc.lights.size() > 1 ) == ( first.sc.lights.size() > 1 ),
                            RG_RESULT_WRONG_FUNCTION_ARGUMENT, "the scenes differ in their light kinds" );
                grp->members.push_back( sc );
                offsets.push_back( offsets.back() + sc->ds.triCount );
            }
            for( RgExtScene_T* sc : grp->members )
            {
                sc->joining = true; // no other group takes it, no destroy frees it
            }
        }
        try
        {
            // the members' triangles, scene after scene, for the batched build
            DeviceBuffer tris;
            tris.Reserve( size_t( offsets.back() ) * 9 * sizeof( float ) );
            for( uint32_t k = 0; k < sceneCount; k++ )
            {
                RG_HIP_CHECK( hipMemcpyAsync( tris.as< float >() + size_t( offsets[ k ] ) * 9, grp->members[ k ]->ds.tris.ptr,
                                              size_t( grp->members[ k ]->ds.triCount ) * 9 * sizeof( float ), hipMemcpyDeviceToDevice, s.setup ) );
            }
            const RgExtBvhBuildParams bp = { 1, s.nodeFormat };
            BuildBvhBatch( grp->bvh, tris.as< float >(), offsets.data(), sceneCount, bp, s.setup );
            RG_HIP_CHECK( hipStreamSynchronize( s.setup ) );
        }
        catch( ... )
        {
            std::lock_guard< std::mutex > lk( s.m );
            for( RgExtScene_T* sc : grp->members )
            {
                sc->joining = false;
            }
            throw;
        }
        grp->bytes = size_t( grp->bvh.nodeCount ) * grp->bvh.NodeStride() + size_t( offsets.back() ) * 48;
        RgExtSceneGroup out = grp.get();
        std::lock_guard< std::mutex > lk( s.m );
        for( uint32_t k = 0; k < sceneCount; k++ )
        {
            grp->members[ k ]->joining    = false;
            grp->members[ k ]->group      = out;
            grp->members[ k ]->groupIndex = k;
        }
        s.sceneBytes += grp->bytes;
        s.groups.push_back( std::move( grp ) );
        *pOut = out;
    } );
}

RgResult rgExtSceneGroupDestroy( RgExtSceneGroup group )
{
    return Guard( [ & ] {
        RG_REQUIRE( group, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "group is null" );
        RgExtServer_T& s = *group->server;
        {
            // later launches trace the scenes' own BVHs
            std::lock_guard< std::mutex > lk( s.m );
            for( RgExtScene_T* sc : group->members )
            {
                sc->group      = nullptr;
                sc->groupIndex = 0;
            }
        }
        s.WaitIdle(); // launches that took the group have finished
        std::unique_ptr< RgExtSceneGroup_T > keep;
        std::lock_guard< std::mutex >        lk( s.m );
        for( auto it = s.groups.begin(); it != s.groups.end(); ++it )
        {
            if( it->get() == group )
            {
                keep = std::move( *it );
                s.groups.erase( it );
                break;
            }
        }
        s.sceneBytes -= group->bytes;
    } );
}

RgResult rgExtSessionCreate( RgExtScene scene, uint32_t width, uint32_t height, RgExtSession* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( scene && pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RG_REQUIRE( scene->built, RG_RESULT_WRONG_FUNCTION_CALL, "the scene is not built" );
        RgExtServer_T& s = *scene->server;
        RG_REQUIRE( width <= s.info.maxRenderWidth && height <= s.info.maxRenderHeight, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "render size exceeds the server's maximum" );
        const size_t bytes = FramebufferSet::BytesFor( width, height, s.storage, true ) + FrameGpu::OwnBytes();
        uint32_t     streamIdx;
        {
            std::lock_guard< std::mutex > lk( s.m );
            if( s.info.memoryBudgetBytes && s.sessionBytes + bytes > s.info.memoryBudgetBytes )
            {
                throw Error( RG_EXT_RESULT_OVER_BUDGET, "the session does not fit the memory budget" );
            }
            s.sessionBytes += bytes; // reserved while the session is set up
            scene->sessions++;
            streamIdx = s.roundRobin++ % uint32_t( s.streams.size() );
        }
        try
        {
            auto ss    = std::make_unique< RgExtSession_T >();
            ss->server = &s;
            ss->scene  = scene;
            ss->width  = width;
            ss->height = height;
            ss->stream = streamIdx;
            ss->bytes  = bytes;
            StreamCtx& st = *s.streams[ streamIdx ];
            ss->fbs.fb.Create( width, height, s.storage, st.planes[ 0 ].get() );
            ss->g.pt.wf = &st.wf;
            if( s.info.flags & RG_EXT_SERVER_READBACK )
            {
                for( void*& img : ss->image )
                {
                    RG_HIP_CHECK( hipHostMalloc( &img, size_t( width ) * height * 4, hipHostMallocDefault ) );
                }
            }
            ss->g.Init( s.setup );
            ss->g.pt.UploadLights( scene->sc, s.setup );
            ss->g.pt.UploadGeomMap( scene->geomMap, s.setup );
            RG_HIP_CHECK( hipStreamSynchronize( s.setup ) );
            RgExtSession out = ss.get();
            std::lock_guard< std::mutex > lk( s.m );
            s.sessions.push_back( std::move( ss ) );
            *pOut = out;
        }
        catch( ... )
        {
            std::lock_guard< std::mutex > lk( s.m );
            s.sessionBytes -= bytes;
            scene->sessions--;
            throw;
        }
    } );
}

RgResult rgExtSessionDestroy( RgExtSession session )
{
    return Guard( [ & ] {
        RG_REQUIRE( session, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "session is null" );
        RgExtServer_T&                    s = *session->server;
        std::unique_ptr< RgExtSession_T > keep;
        {
            std::unique_lock< std::mutex > lk( s.m );
            session->dying = true;
            if( session->pending )
            {
                s.Result( session, session->pending->tag, RG_EXT_FRAME_DROPPED, 0, session->pending->submit, 0.0, NowSeconds() );
                session->pending.reset();
            }
            for( auto it = s.ready.begin(); it != s.ready.end(); ++it )
            {
                if( *it == session )
                {
                    s.ready.erase( it );
                    break;
                }
            }
            s.cvIdle.wait( lk, [ & ] { return session->inflight == 0; } );
            for( auto it = s.sessions.begin(); it != s.sessions.end(); ++it )
            {
                if( it->get() == session )
                {
                    keep = std::move( *it );
                    s.sessions.erase( it );
                    break;
                }
            }
            s.sessionBytes -= session->bytes;
            session->scene->sessions--;
        }
        // results still naming the session stay valid as opaque handles
    } );
}

RgResult rgExtSessionSubmit( RgExtSession session, const RgDrawFrameInfo* pInfo, uint64_t tag )
{
    return Guard( [ & ] {
        RG_REQUIRE( session && pInfo, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        FrameRequest r;
        r.Capture( *pInfo );
        r.tag                            = tag;
        r.submit                         = NowSeconds();
        RgExtServer_T&                   s  = *session->server;
        std::lock_guard< std::mutex > lk( s.m );
        RG_REQUIRE( !session->dying, RG_RESULT_WRONG_FUNCTION_CALL, "the session is being destroyed" );
        if( session->pending )
        {
            s.Result( session, session->pending->tag, RG_EXT_FRAME_DROPPED, 0, session->pending->submit, 0.0, r.submit );
        }
        session->pending = std::move( r );
        if( !session->queued )
        {
            s.ready.push_back( session );
            session->queued = true;
        }
        s.cvWork.notify_one();
    } );
}

RgResult rgExtServerPoll( RgExtServer server, RgExtFrameResult* pResults, uint32_t maxCount, uint32_t* pCount )
{
    return Guard( [ & ] {
        RG_REQUIRE( pCount && ( pResults || maxCount == 0 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        RgExtServer_T&                   s = Srv( server );
        std::lock_guard< std::mutex > lk( s.m );
        uint32_t                      n = 0;
        while( n < maxCount && !s.results.empty() )
        {
            pResults[ n++ ] = s.results.front();
            s.results.pop_front();
        }
        *pCount = n;
    } );
}

RgResult rgExtServerWaitIdle( RgExtServer server )
{
    return Guard( [ & ] { Srv( server ).WaitIdle(); } );
}

RgResult rgExtServerGetStats( RgExtServer server, RgExtServerStats* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pOut is null" );
        RgExtServer_T&                   s = Srv( server );
        std::lock_guard< std::mutex > lk( s.m );
        pOut->sessions      = uint32_t( s.sessions.size() );
        pOut->streams       = uint32_t( s.streams.size() );
        pOut->sessionBytes  = s.sessionBytes;
        pOut->sharedBytes   = s.sharedBytes;
        pOut->sceneBytes    = s.sceneBytes;
        pOut->framesDone    = s.done;
        pOut->framesDropped = s.dropped;
        pOut->framesFailed  = s.failed;
        pOut->now           = NowSeconds();
    } );
}

RgResult rgExtServerSetPolicy( RgExtServer server, const RgExtServerPolicy* pPolicy )
{
    return Guard( [ & ] {
        RgExtServer_T& s = Srv( server );
        RG_REQUIRE( pPolicy, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "pPolicy is null" );
        const RgExtServerPolicy& p = *pPolicy;
        RG_REQUIRE( p.schedule <= RG_EXT_SCHEDULE_NEWEST_FIRST, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "unknown schedule" );
        RG_REQUIRE( p.batchFrames >= 1 && p.batchFrames <= s.batchCapacity, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "batchFrames must be 1 .. the server's batchFrames" );
        RG_REQUIRE( p.framesInFlightPerStream >= 1 && p.framesInFlightPerStream <= 8, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "framesInFlightPerStream must be 1..8" );
        RG_REQUIRE( p.streams >= 1 && p.streams <= s.streams.size(), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "streams must be 1 .. streamCount" );
        RG_REQUIRE( p.maxQueueSeconds >= 0.0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "maxQueueSeconds must not be negative" );
        std::lock_guard< std::mutex > lk( s.m );
        s.info.schedule                = p.schedule;
        s.info.batchFrames             = p.batchFrames;
        s.info.framesInFlightPerStream = p.framesInFlightPerStream;
        s.info.maxQueueSeconds         = p.maxQueueSeconds;
        s.activeStreams                = p.streams;
        s.cvWork.notify_one();
    } );
}

RgResult rgExtSessionGetFramebuffers( RgExtSession session, RgExtFramebuffers* pOut )
{
    return Guard( [ & ] {
        RG_REQUIRE( session && pOut, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "null argument" );
        *pOut = &session->fbs;
    } );
}

} // extern "C"
