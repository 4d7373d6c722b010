#include "rg_frame.h"

#include "rg_denoise.h"
#include "rg_kernels.h"

#include <algorithm>
#include <cmath>

namespace rg
{
namespace
{

// Column-major 4x4 inverse in FP64 (cofactor expansion).
void Inverse4( const float* m, float* out )
{
    double a[ 16 ], inv[ 16 ];
    for( int i = 0; i < 16; i++ )
    {
        a[ i ] = m[ i ];
    }
    inv[ 0 ]  = a[ 5 ] * a[ 10 ] * a[ 15 ] - a[ 5 ] * a[ 11 ] * a[ 14 ] - a[ 9 ] * a[ 6 ] * a[ 15 ] + a[ 9 ] * a[ 7 ] * a[ 14 ] + a[ 13 ] * a[ 6 ] * a[ 11 ] - a[ 13 ] * a[ 7 ] * a[ 10 ];
    inv[ 4 ]  = -a[ 4 ] * a[ 10 ] * a[ 15 ] + a[ 4 ] * a[ 11 ] * a[ 14 ] + a[ 8 ] * a[ 6 ] * a[ 15 ] - a[ 8 ] * a[ 7 ] * a[ 14 ] - a[ 12 ] * a[ 6 ] * a[ 11 ] + a[ 12 ] * a[ 7 ] * a[ 10 ];
    inv[ 8 ]  = a[ 4 ] * a[ 9 ] * a[ 15 ] - a[ 4 ] * a[ 11 ] * a[ 13 ] - a[ 8 ] * a[ 5 ] * a[ 15 ] + a[ 8 ] * a[ 7 ] * a[ 13 ] + a[ 12 ] * a[ 5 ] * a[ 11 ] - a[ 12 ] * a[ 7 ] * a[ 9 ];
    inv[ 12 ] = -a[ 4 ] * a[ 9 ] * a[ 14 ] + a[ 4 ] * a[ 10 ] * a[ 13 ] + a[ 8 ] * a[ 5 ] * a[ 14 ] - a[ 8 ] * a[ 6 ] * a[ 13 ] - a[ 12 ] * a[ 5 ] * a[ 10 ] + a[ 12 ] * a[ 6 ] * a[ 9 ];
    inv[ 1 ]  = -a[ 1 ] * a[ 10 ] * a[ 15 ] + a[ 1 ] * a[ 11 ] * a[ 14 ] + a[ 9 ] * a[ 2 ] * a[ 15 ] - a[ 9 ] * a[ 3 ] * a[ 14 ] - a[ 13 ] * a[ 2 ] * a[ 11 ] + a[ 13 ] * a[ 3 ] * a[ 10 ];
    inv[ 5 ]  = a[ 0 ] * a[ 10 ] * a[ 15 ] - a[ 0 ] * a[ 11 ] * a[ 14 ] - a[ 8 ] * a[ 2 ] * a[ 15 ] + a[ 8 ] * a[ 3 ] * a[ 14 ] + a[ 12 ] * a[ 2 ] * a[ 11 ] - a[ 12 ] * a[ 3 ] * a[ 10 ];
    inv[ 9 ]  = -a[ 0 ] * a[ 9 ] * a[ 15 ] + a[ 0 ] * a[ 11 ] * a[ 13 ] + a[ 8 ] * a[ 1 ] * a[ 15 ] - a[ 8 ] * a[ 3 ] * a[ 13 ] - a[ 12 ] * a[ 1 ] * a[ 11 ] + a[ 12 ] * a[ 3 ] * a[ 9 ];
    inv[ 13 ] = a[ 0 ] * a[ 9 ] * a[ 14 ] - a[ 0 ] * a[ 10 ] * a[ 13 ] - a[ 8 ] * a[ 1 ] * a[ 14 ] + a[ 8 ] * a[ 2 ] * a[ 13 ] + a[ 12 ] * a[ 1 ] * a[ 10 ] - a[ 12 ] * a[ 2 ] * a[ 9 ];
    inv[ 2 ]  = a[ 1 ] * a[ 6 ] * a[ 15 ] - a[ 1 ] * a[ 7 ] * a[ 14 ] - a[ 5 ] * a[ 2 ] * a[ 15 ] + a[ 5 ] * a[ 3 ] * a[ 14 ] + a[ 13 ] * a[ 2 ] * a[ 7 ] - a[ 13 ] * a[ 3 ] * a[ 6 ];
    inv[ 6 ]  = -a[ 0 ] * a[ 6 ] * a[ 15 ] + a[ 0 ] * a[ 7 ] * a[ 14 ] + a[ 4 ] * a[ 2 ] * a[ 15 ] - a[ 4 ] * a[ 3 ] * a[ 14 ] - a[ 12 ] * a[ 2 ] * a[ 7 ] + a[ 12 ] * a[ 3 ] * a[ 6 ];
    inv[ 10 ] = a[ 0 ] * a[ 5 ] * a[ 15 ] - a[ 0 ] * a[ 7 ] * a[ 13 ] - a[ 4 ] * a[ 1 ] * a[ 15 ] + a[ 4 ] * a[ 3 ] * a[ 13 ] + a[ 12 ] * a[ 1 ] * a[ 7 ] - a[ 12 ] * a[ 3 ] * a[ 5 ];
    inv[ 14 ] = -a[ 0 ] * a[ 5 ] * a[ 14 ] + a[ 0 ] * a[ 6 ] * a[ 13 ] + a[ 4 ] * a[ 1 ] * a[ 14 ] - a[ 4 ] * a[ 2 ] * a[ 13 ] - a[ 12 ] * a[ 1 ] * a[ 6 ] + a[ 12 ] * a[ 2 ] * a[ 5 ];
    inv[ 3 ]  = -a[ 1 ] * a[ 6 ] * a[ 11 ] + a[ 1 ] * a[ 7 ] * a[ 10 ] + a[ 5 ] * a[ 2 ] * a[ 11 ] - a[ 5 ] * a[ 3 ] * a[ 10 ] - a[ 9 ] * a[ 2 ] * a[ 7 ] + a[ 9 ] * a[ 3 ] * a[ 6 ];
    inv[ 7 ]  = a[ 0 ] * a[ 6 ] * a[ 11 ] - a[ 0 ] * a[ 7 ] * a[ 10 ] - a[ 4 ] * a[ 2 ] * a[ 11 ] + a[ 4 ] * a[ 3 ] * a[ 10 ] + a[ 8 ] * a[ 2 ] * a[ 7 ] - a[ 8 ] * a[ 3 ] * a[ 6 ];
    inv[ 11 ] = -a[ 0 ] * a[ 5 ] * a[ 11 ] + a[ 0 ] * a[ 7 ] * a[ 9 ] + a[ 4 ] * a[ 1 ] * a[ 11 ] - a[ 4 ] * a[ 3 ] * a[ 9 ] - a[ 8 ] * a[ 1 ] * a[ 7 ] + a[ 8 ] * a[ 3 ] * a[ 5 ];
    inv[ 15 ] = a[ 0 ] * a[ 5 ] * a[ 10 ] - a[ 0 ] * a[ 6 ] * a[ 9 ] - a[ 4 ] * a[ 1 ] * a[ 10 ] + a[ 4 ] * a[ 2 ] * a[ 9 ] + a[ 8 ] * a[ 1 ] * a[ 6 ] - a[ 8 ] * a[ 2 ] * a[ 5 ];
    double det = a[ 0 ] * inv[ 0 ] + a[ 1 ] * inv[ 4 ] + a[ 2 ] * inv[ 8 ] + a[ 3 ] * inv[ 12 ];
    det        = det != 0.0 ? 1.0 / det : 0.0;
    for( int i = 0; i < 16; i++ )
    {
        out[ i ] = float( inv[ i ] * det );
    }
}

// Matrix::MakeProjectionMatrix (Source/Matrix.cpp:380-395)
void MakeProjection( float* m, float aspect, float fovY, float zn, float zf )
{
    const float t = tanf( fovY * 0.5f );
    std::memset( m, 0, 16 * sizeof( float ) );
    m[ 0 ]  = 1.0f / ( aspect * t );
    m[ 5 ]  = -1.0f / t;
    m[ 10 ] = zf / ( zn - zf );
    m[ 11 ] = -1.0f;
    m[ 14 ] = ( zf * zn ) / ( zn - zf );
}

} // namespace

CameraState::CameraState()
{
    rgExtFrameParamsDefaults( 1280, 720, &params );
}

// FillUniform (Source/VulkanDevice.cpp:142-482), hot-path fields only.
void FillFrameParams( CameraState& cs, const RgDrawFrameInfo& di, uint32_t w, uint32_t h, uint32_t lightArraySize, uint32_t prevLightArraySize,
                      bool hasDirectional, bool cullBackFaces, bool allowSkyGeometry, uint32_t ignoreFpvLightIndex )
{
    RgExtFrameParams& p = cs.params;
    std::memcpy( p.viewPrev, p.view, sizeof( p.view ) );
    std::memcpy( p.projectionPrev, p.projection, sizeof( p.projection ) );
    std::memcpy( p.cameraPositionPrev, p.cameraPosition, sizeof( p.cameraPosition ) );
    std::memcpy( p.view, di.view, sizeof( p.view ) );
    MakeProjection( p.projection, float( w ) / float( h ), di.fovYRadians, di.cameraNear, di.cameraFar );
    Inverse4( p.view, p.invView );
    Inverse4( p.projection, p.invProjection );
    p.cameraPosition[ 0 ] = p.invView[ 12 ];
    p.cameraPosition[ 1 ] = p.invView[ 13 ];
    p.cameraPosition[ 2 ] = p.invView[ 14 ];
    p.cameraPosition[ 3 ] = 0.0f;
    if( cs.firstFrame )
    {
        std::memcpy( p.viewPrev, p.view, sizeof( p.view ) );
        std::memcpy( p.projectionPrev, p.projection, sizeof( p.projection ) );
        std::memcpy( p.cameraPositionPrev, p.cameraPosition, sizeof( p.cameraPosition ) );
    }
    p.renderWidth  = w;
    p.renderHeight = h;
    p.frameId      = cs.frameId;
    p.timeDelta    = float( std::max( di.currentTime - cs.prevTime, 0.001 ) );
    cs.prevTime    = di.currentTime;
    p.jitterX = p.jitterY = 0.0f;

    const auto* tm      = FindParams< RgDrawFrameTonemappingParams >( di, RG_STRUCTURE_TYPE_TONEMAPPING );
    const float ev100Min = tm ? tm->ev100Min : 0.0f, ev100Max = tm ? tm->ev100Max : 10.0f;
    p.stopEyeAdaptation  = tm ? tm->disableEyeAdaptation : 0;
    p.minLogLuminance    = std::log2( std::exp2( ev100Min ) * 12.5f / 100.0f );
    p.maxLogLuminance    = std::log2( std::exp2( ev100Max ) * 12.5f / 100.0f );

    p.lightCount             = lightArraySize - 1u;
    p.lightCountPrev         = prevLightArraySize > 0 ? prevLightArraySize - 1u : 0u;
    p.directionalLightExists = hasDirectional ? 1u : 0u;

    const auto* sky = FindParams< RgDrawFrameSkyParams >( di, RG_STRUCTURE_TYPE_SKY );
    if( sky )
    {
        for( int i = 0; i < 3; i++ )
        {
            p.skyColorDefault[ i ] = sky->skyColorDefault.data[ i ];
        }
        p.skyColorMultiplier = std::max( 0.0f, sky->skyColorMultiplier );
        p.skyColorSaturation = std::max( 0.0f, sky->skyColorSaturation );
    }
    const auto* il = FindParams< RgDrawFrameIlluminationParams >( di, RG_STRUCTURE_TYPE_ILLUMINATION );
    p.maxBounceShadowsLights   = il ? il->maxBounceShadows : 2u;
    p.polyLightSpotlightFactor = il ? std::max( 0.0f, il->polygonalLightSpotlightFactor ) : 2.0f;
    p.indirSecondBounce        = il ? !!il->enableSecondBounceForIndirect : 1u;
    p.cellWorldSize            = il ? std::max( il->cellWorldSize, 0.001f ) : 1.0f;
    p.gradientMultDiffuse      = il ? std::clamp( il->directDiffuseSensitivityToChange, 0.0f, 1.0f ) : 0.5f;
    p.gradientMultIndirect     = il ? std::clamp( il->indirectDiffuseSensitivityToChange, 0.0f, 1.0f ) : 0.2f;
    p.gradientMultSpecular     = il ? std::clamp( il->specularSensitivityToChange, 0.0f, 1.0f ) : 0.5f;
    const auto* tx             = FindParams< RgDrawFrameTexturesParams >( di, RG_STRUCTURE_TYPE_TEXTURES );
    p.minRoughness             = tx ? std::clamp( tx->minRoughness, 0.0f, 1.0f ) : 0.0f;
    p.emissionMapBoost         = tx ? std::max( tx->emissionMapBoost, 0.0f ) : 100.0f;
    p.normalMapStrength        = tx ? tx->normalMapStrength : 1.0f;
    p.emissionMaxScreenColor   = tx ? std::max( tx->emissionMaxScreenColor, 0.0f ) : 1.5f;

    p.rayCullBackFaces  = cullBackFaces ? 1u : 0u;
    p.rayLength         = std::clamp( di.rayLength, 0.1f, 10000.0f );
    p.primaryRayMinDist = std::clamp( di.cameraNear, 0.001f, p.rayLength );
    p.antiFireflyEnabled = 1;

    const auto* ext = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
    p.atrousIterations        = ext && ext->atrousIterations ? ext->atrousIterations : 4u;
    p.restirInitialCandidates = ext && ext->restirInitialCandidates ? ext->restirInitialCandidates : 8u;
    p.restirTemporalSamples   = ext && ext->restirTemporalSamples != 0xFFFFFFFFu ? ext->restirTemporalSamples : 1u;
    p.restirSpatialSamples    = ext && ext->restirSpatialSamples != 0xFFFFFFFFu ? ext->restirSpatialSamples : 8u;
    RG_REQUIRE( p.atrousIterations == 4 || p.atrousIterations == 5, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "atrousIterations must be 4 or 5" );

    // Source/VulkanDevice.cpp:306-339
    const auto* rr           = FindParams< RgDrawFrameReflectRefractParams >( di, RG_STRUCTURE_TYPE_REFLECTREFRACT );
    const uint32_t media     = rr ? uint32_t( rr->typeOfMediaAroundCamera ) : uint32_t( RG_MEDIA_TYPE_VACUUM );
    p.cameraMediaType        = media <= uint32_t( RG_MEDIA_TYPE_ACID ) ? media : uint32_t( RG_MEDIA_TYPE_VACUUM );
    p.reflectRefractMaxDepth = std::min( 16u, rr ? rr->maxReflectRefractDepth : 2u );
    p.indexOfRefractionGlass = std::max( 0.0f, rr ? rr->indexOfRefractionGlass : 1.52f );
    p.indexOfRefractionWater = std::max( 0.0f, rr ? rr->indexOfRefractionWater : 1.33f );
    p.thinMediaWidth         = std::max( 0.0f, rr ? rr->thinMediaWidth : 0.1f );
    p.cameraRayConeSpreadAngle = std::atan( ( 2.0f * std::tan( di.fovYRadians * 0.5f ) ) / float( h ) );
    p.waterWaveStrength        = rr ? rr->waterWaveNormalStrength : 1.0f;
    const float water[ 3 ] = { 0.3f, 0.73f, 0.63f }, acid[ 3 ] = { 0.0f, 0.66f, 0.55f };
    for( int i = 0; i < 3; i++ )
    {
        p.waterColorAndDensity[ i ] = rr ? rr->waterColor.data[ i ] : water[ i ];
        p.acidColorAndDensity[ i ]  = rr ? rr->acidColor.data[ i ] : acid[ i ];
    }
    p.waterColorAndDensity[ 3 ] = 0.0f;
    p.acidColorAndDensity[ 3 ]  = std::max( 0.0f, rr ? rr->acidDensity : 10.0f );
    // SceneImportExport's worldUp: SafeNormalize(worldUp, {0,1,0}) (Source/Scene.cpp:385)
    const double ux = cs.worldUp[ 0 ], uy = cs.worldUp[ 1 ], uz = cs.worldUp[ 2 ];
    const double ul = std::sqrt( ux * ux + uy * uy + uz * uz );
    p.worldUp[ 0 ] = ul > 1e-12 ? float( ux / ul ) : 0.0f;
    p.worldUp[ 1 ] = ul > 1e-12 ? float( uy / ul ) : 1.0f;
    p.worldUp[ 2 ] = ul > 1e-12 ? float( uz / ul ) : 0.0f;
    p.worldUp[ 3 ] = 0.0f;

    // ray cull masks (Source/VulkanDevice.cpp:344-393)
    uint32_t world = 0;
    if( di.rayCullMaskWorld & RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT )
    {
        world |= IMASK_WORLD_0;
    }
    if( di.rayCullMaskWorld & RG_DRAW_FRAME_RAY_CULL_WORLD_1_BIT )
    {
        world |= IMASK_WORLD_1;
    }
    if( di.rayCullMaskWorld & RG_DRAW_FRAME_RAY_CULL_WORLD_2_BIT )
    {
        RG_REQUIRE( !allowSkyGeometry, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "RG_DRAW_FRAME_RAY_CULL_WORLD_2_BIT cannot be used, as RgInstanceCreateInfo::allowGeometryWithSkyFlag was true" );
        world |= IMASK_WORLD_2;
    }
    if( di.rayCullMaskWorld & RG_DRAW_FRAME_RAY_CULL_SKY_BIT )
    {
        RG_REQUIRE( allowSkyGeometry, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "RG_DRAW_FRAME_RAY_CULL_SKY_BIT cannot be used, as RgInstanceCreateInfo::allowGeometryWithSkyFlag was false" );
        world |= IMASK_WORLD_2;
    }
    p.rayCullMaskWorld           = world;
    p.rayCullMaskWorldShadow     = allowSkyGeometry ? ( world & ~IMASK_WORLD_2 ) : world;
    p.lightIndexIgnoreFPVShadows = ignoreFpvLightIndex;
    p.skyGeometry                = allowSkyGeometry ? 1u : 0u;
}

void RenderSizeOf( const RgDrawFrameInfo& di, uint32_t* w, uint32_t* h )
{
    *w      = 1280;
    *h      = 720;
    auto rr = FindParams< RgDrawFrameRenderResolutionParams >( di, RG_STRUCTURE_TYPE_RENDER_RESOLUTION );
    if( rr && rr->resolutionMode == RG_RENDER_RESOLUTION_MODE_CUSTOM && rr->customRenderSize.width && rr->customRenderSize.height )
    {
        *w = rr->customRenderSize.width;
        *h = rr->customRenderSize.height;
    }
    auto ext = FindParams< RgExtDrawFramePathParams >( di, RG_EXT_STRUCTURE_TYPE_PATH_PARAMS );
    if( ext && ext->renderWidth && ext->renderHeight )
    {
        *w = ext->renderWidth;
        *h = ext->renderHeight;
    }
    *w &= ~1u; // render width must be even (VulkanDevice.cpp:186-187)
    RG_REQUIRE( *w >= 2 && *h >= 1, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render size too small" );
}

void FrameGpu::Init( hipStream_t s )
{
    hist.Reserve( 256 * 4 );
    histCopy.Reserve( 256 * 4 );
    avgLum.Reserve( 4 );
    RG_HIP_CHECK( hipMemsetAsync( hist.ptr, 0, 256 * 4, s ) );
    RG_HIP_CHECK( hipMemsetAsync( histCopy.ptr, 0, 256 * 4, s ) );
    RG_HIP_CHECK( hipMemsetAsync( avgLum.ptr, 0, 4, s ) );
    pt.Init( s );
}

namespace
{
// A-SVGF, eye adaptation and the final compose of one traced frame.
void FinishFrame( FrameGpu& g, const RgExtFrameParams& p, FramebufferSet& fb, const RgDrawFrameInfo& di, hipStream_t s, hipEvent_t evDenoised )
{
    const uint32_t w = p.renderWidth, h = p.renderHeight;
    Denoise( fb, p, s );
    if( evDenoised )
    {
        RG_HIP_CHECK( hipEventRecord( evDenoised, s ) );
    }
    LuminanceHistogram( reinterpret_cast< const float* >( fb.Cur( RG_EXT_PLANE_HISTOGRAM_INPUT ) ), w * h, p.minLogLuminance, p.maxLogLuminance,
                        g.hist.as< uint32_t >(), s );
    RG_HIP_CHECK( hipMemcpyAsync( g.histCopy.ptr, g.hist.ptr, 256 * 4, hipMemcpyDeviceToDevice, s ) );
    LuminanceAverage( g.hist.as< uint32_t >(), w * h, p.minLogLuminance, p.maxLogLuminance, p.timeDelta, p.stopEyeAdaptation != 0,
                      g.avgLum.as< float >(), s );
    // ImageComposition::Finalize (Source/ImageComposition.cpp:94-156)
    const auto* tm       = FindParams< RgDrawFrameTonemappingParams >( di, RG_STRUCTURE_TYPE_TONEMAPPING );
    float       sat[ 3 ] = { 0.0f, 0.0f, 0.0f }, ct[ 3 ] = { 1.0f, 1.0f, 1.0f }; // DrawFrameInfo.h:104-105
    if( tm )
    {
        std::memcpy( sat, tm->saturation.data, sizeof( sat ) );
        std::memcpy( ct, tm->crosstalk.data, sizeof( ct ) );
    }
    ComposeFinal( fb, MakeLpmConsts( sat, ct ), p.emissionMaxScreenColor, g.avgLum.as< float >(), s );
}
} // namespace

uint64_t EncodeFrame( FrameGpu& g, TraceContext& tc, const RgExtFrameParams& p, const BvhObject& bvh, const DeviceScene& scene,
                      FramebufferSet& fb, const RgDrawFrameInfo& di, hipStream_t s, hipEvent_t evTraced, hipEvent_t evDenoised )
{
    const uint32_t flags = p.rayCullBackFaces ? RG_EXT_TRACE_CULL_FRONT_FACE : 0u;
    const uint64_t rays  = g.pt.Trace( p, bvh, tc, scene, fb, flags, s, nullptr );
    if( evTraced )
    {
        RG_HIP_CHECK( hipEventRecord( evTraced, s ) );
    }
    FinishFrame( g, p, fb, di, s, evDenoised );
    return rays;
}

uint64_t EncodeFrames( const EncodeSlot* slots, uint32_t B, TraceContext& tc, const BvhObject& bvh, const DeviceScene* scene, Wavefront& wf,
                       hipStream_t s )
{
    RG_REQUIRE( B >= 1 && B <= MAX_TRACE_SLOTS, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "frames per batch out of range" );
    TraceSlot ts[ MAX_TRACE_SLOTS ];
    for( uint32_t k = 0; k < B; k++ )
    {
        ts[ k ] = TraceSlot{ &slots[ k ].g->pt, slots[ k ].p, slots[ k ].fb, slots[ k ].scene, slots[ k ].group };
    }
    const uint32_t flags = slots[ 0 ].p->rayCullBackFaces ? RG_EXT_TRACE_CULL_FRONT_FACE : 0u;
    const uint64_t rays  = TraceFrames( ts, B, bvh, tc, scene, wf, flags, s );
    for( uint32_t k = 0; k < B; k++ )
    {
        FinishFrame( *slots[ k ].g, *slots[ k ].p, *slots[ k ].fb, *slots[ k ].di, s, nullptr );
    }
    return rays;
}

} // namespace rg
