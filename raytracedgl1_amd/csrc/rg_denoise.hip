// A-SVGF on gfx950: gradient a-trous, temporal accumulation, anti-firefly,
// variance estimation, edge-avoiding a-trous and final composition.
//
// Restates the reference compute shaders (CmASVGFGradientAtrous.comp,
// CmSVGFTemporalAccumulation.comp, CmAntiFirefly.comp,
// CmSVGFEstimateVariance.comp, CmSVGFAtrous_Iter0.comp, CmSVGFAtrous.comp)
// in the pass order of Denoiser::Denoise (Source/Denoiser.cpp:64-255).
//
// Layout: 256-thread blocks of 32 x 8 pixels; a wave covers two rows of 32
// pixels, so each wave's loads of a 16-byte plane are two 512-byte
// contiguous runs.  The temporal pass decodes each pixel's normal once and
// writes it with the depth to the guide plane (16 bytes per pixel); the
// variance and a-trous passes read one guide value per tap instead of
// decoding the normal (two sincospi) at every tap.  Iteration-0 passes
// stage their footprint (tile + halo) in LDS once per block.  A fetch
// outside the image returns zeros, like texelFetch with robust image access.
#include "rg_denoise.h"
#include "rg_device_math.h"

namespace rg
{
namespace
{

constexpr int   BX = 32, BY = 8, THREADS = BX * BY;
constexpr float MAX_RAY_LENGTH = 10000.0f; // ShaderCommonC.h:197
constexpr int   STRATA = 3;                // COMPUTE_ASVGF_STRATA_SIZE

struct Dims
{
    int W, H, GW, GH;
};

__device__ __forceinline__ float4 Ld4( const void* base, int w, int h, int x, int y )
{
    if( x < 0 || y < 0 || x >= w || y >= h )
    {
        return make_float4( 0.f, 0.f, 0.f, 0.f );
    }
    return reinterpret_cast< const float4* >( base )[ size_t( y ) * w + x ];
}
__device__ __forceinline__ float2 Ld2( const void* base, int w, int h, int x, int y )
{
    if( x < 0 || y < 0 || x >= w || y >= h )
    {
        return make_float2( 0.f, 0.f );
    }
    return reinterpret_cast< const float2* >( base )[ size_t( y ) * w + x ];
}
__device__ __forceinline__ float Ld1( const void* base, int w, int h, int x, int y )
{
    if( x < 0 || y < 0 || x >= w || y >= h )
    {
        return 0.f;
    }
    return reinterpret_cast< const float* >( base )[ size_t( y ) * w + x ];
}
__device__ __forceinline__ uint32_t LdU( const void* base, int w, int h, int x, int y )
{
    if( x < 0 || y < 0 || x >= w || y >= h )
    {
        return 0u;
    }
    return reinterpret_cast< const uint32_t* >( base )[ size_t( y ) * w + x ];
}
__device__ __forceinline__ void St4( void* base, int w, int x, int y, float4 v )
{
    reinterpret_cast< float4* >( base )[ size_t( y ) * w + x ] = v;
}
__device__ __forceinline__ void St2( void* base, int w, int x, int y, float2 v )
{
    reinterpret_cast< float2* >( base )[ size_t( y ) * w + x ] = v;
}
__device__ __forceinline__ void St1( void* base, int w, int x, int y, float v )
{
    reinterpret_cast< float* >( base )[ size_t( y ) * w + x ] = v;
}
// Loads and stores of the planes that RG_EXT_STORAGE_FP16 keeps as half
// (IsHalfPlane): F16 selects the 2-byte form.  Stores round to nearest and
// clamp to the largest finite half, so an overflow cannot become inf.
typedef _Float16 half4v __attribute__( ( ext_vector_type( 4 ) ) );
typedef _Float16 half2v __attribute__( ( ext_vector_type( 2 ) ) );
constexpr float HALF_MAX = 65504.0f;
__device__ __forceinline__ _Float16 ToHalf( float v )
{
    return _Float16( fminf( fmaxf( v, -HALF_MAX ), HALF_MAX ) );
}
template< bool F16 >
__device__ __forceinline__ float4 LdP4( const void* base, int w, int h, int x, int y )
{
    if constexpr( !F16 )
    {
        return Ld4( base, w, h, x, y );
    }
    else
    {
        if( x < 0 || y < 0 || x >= w || y >= h )
        {
            return make_float4( 0.f, 0.f, 0.f, 0.f );
        }
        const half4v v = reinterpret_cast< const half4v* >( base )[ size_t( y ) * w + x ];
        return make_float4( float( v.x ), float( v.y ), float( v.z ), float( v.w ) );
    }
}
template< bool F16 >
__device__ __forceinline__ float2 LdP2( const void* base, int w, int h, int x, int y )
{
    if constexpr( !F16 )
    {
        return Ld2( base, w, h, x, y );
    }
    else
    {
        if( x < 0 || y < 0 || x >= w || y >= h )
        {
            return make_float2( 0.f, 0.f );
        }
        const half2v v = reinterpret_cast< const half2v* >( base )[ size_t( y ) * w + x ];
        return make_float2( float( v.x ), float( v.y ) );
    }
}
template< bool F16 >
__device__ __forceinline__ float LdP1( const void* base, int w, int h, int x, int y )
{
    if constexpr( !F16 )
    {
        return Ld1( base, w, h, x, y );
    }
    else
    {
        if( x < 0 || y < 0 || x >= w || y >= h )
        {
            return 0.f;
        }
        return float( reinterpret_cast< const _Float16* >( base )[ size_t( y ) * w + x ] );
    }
}
template< bool F16 >
__device__ __forceinline__ void StP4( void* base, int w, int x, int y, float4 v )
{
    if constexpr( !F16 )
    {
        St4( base, w, x, y, v );
    }
    else
    {
        half4v h;
        h.x = ToHalf( v.x ), h.y = ToHalf( v.y ), h.z = ToHalf( v.z ), h.w = ToHalf( v.w );
        reinterpret_cast< half4v* >( base )[ size_t( y ) * w + x ] = h;
    }
}
template< bool F16 >
__device__ __forceinline__ void StP2( void* base, int w, int x, int y, float2 v )
{
    if constexpr( !F16 )
    {
        St2( base, w, x, y, v );
    }
    else
    {
        half2v h;
        h.x = ToHalf( v.x ), h.y = ToHalf( v.y );
        reinterpret_cast< half2v* >( base )[ size_t( y ) * w + x ] = h;
    }
}
template< bool F16 >
__device__ __forceinline__ void StP1( void* base, int w, int x, int y, float v )
{
    if constexpr( !F16 )
    {
        St1( base, w, x, y, v );
    }
    else
    {
        reinterpret_cast< _Float16* >( base )[ size_t( y ) * w + x ] = ToHalf( v );
    }
}
__device__ __forceinline__ v3 xyz( float4 a )
{
    return mk3( a.x, a.y, a.z );
}
__device__ __forceinline__ float4 f4( v3 a, float w )
{
    return make_float4( a.x, a.y, a.z, w );
}

// decodeNormal, Source/Shaders/Utils.h:126-140: phi = j pi / 65535,
// theta = k 2 pi / 65536, evaluated with sincospi.
__device__ __forceinline__ v3 DecodeNormal( uint32_t p )
{
    float sp, cp, st, ct;
    sincospif( unitQuotient< 65535u >( p >> 16 ), &sp, &cp );
    sincospif( float( p & 0xFFFFu ) / 32768.0f, &st, &ct );
    return mk3( sp * ct, sp * st, cp );
}

// The denoise guide of a pixel: its decoded normal and linear depth, as
// DecodeNormal(NORMAL) and DEPTH_WORLD would give them; outside the image
// the values of a zero fetch of both (decoded normal code 0, depth 0).
__device__ __forceinline__ void LdGuide( const FbPtrs& P, int w, int h, int x, int y, v3& n, float& depth )
{
    if( x < 0 || y < 0 || x >= w || y >= h )
    {
        n     = DecodeNormal( 0u );
        depth = 0.0f;
        return;
    }
    const float4 g = reinterpret_cast< const float4* >( P.cur[ RG_EXT_PLANE_DENOISE_GUIDE ] )[ size_t( y ) * w + x ];
    n              = mk3( g.x, g.y, g.z );
    depth          = g.w;
}

// Transcendentals of filter weights and accumulation speeds on the hardware
// units: v_exp_f32 / v_log_f32 (base 2) and v_rcp_f32, each within about
// one ulp, instead of the libm sequences (powf alone is some sixty
// instructions).  These values only weight normalised averages or set
// blend factors, so their relative error (about 1e-6 over the arguments
// the passes produce) moves the denoised radiance by no more than that,
// against its 1e-3 bound.  Comparisons that choose a branch (reprojection
// tests, history thresholds) keep IEEE division.
__device__ __forceinline__ float FastExp( float x )
{
    return __builtin_amdgcn_exp2f( x * 1.44269504088896341f );
}
__device__ __forceinline__ float FastExp2( float x )
{
    return __builtin_amdgcn_exp2f( x );
}
// x^y for x >= 0 and y > 0 (pow(0, y) = 0 through log2(0) = -inf)
__device__ __forceinline__ float FastPow( float x, float y )
{
    return __builtin_amdgcn_exp2f( y * __builtin_amdgcn_logf( x ) );
}
__device__ __forceinline__ float FastLn( float x )
{
    return __builtin_amdgcn_logf( x ) * 0.693147180559945309f;
}
__device__ __forceinline__ float Pow5( float x )
{
    const float x2 = x * x;
    return x2 * x2 * x;
}
__device__ __forceinline__ float Pow10( float x )
{
    const float x5 = Pow5( x );
    return x5 * x5;
}
__device__ __forceinline__ float Pow128( float x )
{
#pragma unroll
    for( int i = 0; i < 7; i++ )
    {
        x *= x;
    }
    return x;
}

__device__ __forceinline__ bool TestDepth( float z, float zPrev, float zMotion )
{
    return fabsf( z - zPrev + zMotion ) / fabsf( z ) < 0.1f;
}
__device__ __forceinline__ bool TestNormal( v3 n, v3 nPrev )
{
    return dot( n, nPrev ) > 0.95f;
}

// getPrevScreenPos (ShaderCommonGLSLFunc.h:290-297): the checkerboard half is
// W/2 wide, so motion in UV units is scaled by W/2.
__device__ __forceinline__ void PrevScreenPos( float mx, float my, int x, int y, int W, int H, float* px, float* py )
{
    const float sx = float( W ) / 2.0f, sy = float( H );
    *px            = ( ( float( x ) + 0.5f ) * ( 1.0f / sx ) + mx ) * sx;
    *py            = ( ( float( y ) + 0.5f ) * ( 1.0f / sy ) + my ) * sy;
}

__device__ __forceinline__ void MulCM( const float* m, v3 v, float w, float out[ 4 ] )
{
#pragma unroll
    for( int r = 0; r < 4; r++ )
    {
        out[ r ] = m[ 0 * 4 + r ] * v.x + m[ 1 * 4 + r ] * v.y + m[ 2 * 4 + r ] * v.z + m[ 3 * 4 + r ] * w;
    }
}

__device__ __forceinline__ float SpecMinAccumFrames( const RgExtFrameParams& g )
{
    return clampf( 0.2f * safePositiveRcp( g.timeDelta ), 0.0f, 128.0f );
}

// CmSVGFTemporalAccumulation.comp:40-86: 60 tan of the angle between the
// surface point's view directions of this and the previous frame.  The sine
// is |v x vPrev| rather than sqrt(1 - cos^2): the same value for unit
// vectors, but sqrt(1 - cos^2) cancels catastrophically in FP32 at the small
// angles of a slowly moving camera (cos within an ulp of 1 leaves no digits
// of the angle), where the cross product keeps every digit.  At cos <= 0
// the clamped cosine makes the sine 1, as in the shader.
__device__ float Parallax( const RgExtFrameParams& g, v3 pos, v3 posPrev )
{
    const v3    camDelta = mk3( g.cameraPositionPrev[ 0 ] - g.cameraPosition[ 0 ],
                             g.cameraPositionPrev[ 1 ] - g.cameraPosition[ 1 ],
                             g.cameraPositionPrev[ 2 ] - g.cameraPosition[ 2 ] );
    const v3    v        = normalize( pos );
    const v3    vPrev    = normalize( posPrev - camDelta );
    const float c        = dot( v, vPrev );
    const float cosa     = saturate( c );
    const float sina     = c > 0.0f ? fminf( length( cross( v, vPrev ) ), 1.0f ) : 1.0f;
    return sina / fmaxf( cosa, 1e-6f ) * 60.0f;
}
__device__ float SpecAccumSpeed( const RgExtFrameParams& g, float Amax, float roughness, float NoV, float parallax )
{
    const float a                   = FastPow( saturate( 1.0f - NoV ), 0.66f );
    const float b                   = 1.1f + roughness * roughness;
    const float parallaxSensitivity = ( b + a ) / ( b - a );
    const float powerScale          = 1.0f + parallax * parallaxSensitivity;
    float       f                   = 1.0f - FastExp2( -200.0f * roughness * roughness );
    f *= FastPow( saturate( roughness ), 0.5f * powerScale );
    return clampf( 256.0f * f, SpecMinAccumFrames( g ), Amax );
}
__device__ float SpecularDominantFactor( float nv, float roughness )
{
    const float a = 0.298475f * FastLn( 39.4115f - 39.0029f * roughness );
    const float f = FastPow( clampf( 1.0f - nv, 0.0f, 1.0f ), 10.8649f ) * ( 1.0f - a ) + a;
    return clampf( f, 0.0f, 1.0f );
}

// ------------------------------------------------------ gradient a-trous

__global__ __launch_bounds__( THREADS ) void k_grad_atrous( FbPtrs P, Dims d, int iter )
{
    const int gx = blockIdx.x * BX + threadIdx.x % BX;
    const int gy = blockIdx.y * BY + threadIdx.x / BX;
    if( gx >= d.GW || gy >= d.GH )
    {
        return;
    }
    const void* src  = P.cur[ iter % 2 == 0 ? RG_EXT_PLANE_DIS_PING_GRADIENT : RG_EXT_PLANE_DIS_PONG_GRADIENT ];
    void*       dst  = P.cur[ iter % 2 == 0 ? RG_EXT_PLANE_DIS_PONG_GRADIENT : RG_EXT_PLANE_DIS_PING_GRADIENT ];
    const int   step = 1 << iter;
    int         lo, hi;
    cbRenderArea( gx * STRATA, d.W, &lo, &hi );
    lo /= STRATA;
    hi /= STRATA;
    const int gh = d.H / STRATA;
    v3        acc  = mk3( 0, 0, 0 );
    float     wsum = 0.0f;
    for( int yy = -1; yy <= 1; yy++ )
    {
        for( int xx = -1; xx <= 1; xx++ )
        {
            const int   qx = gx + xx * step, qy = gy + yy * step;
            const float k  = ( xx == 0 ? 1.0f : 0.5f ) * ( yy == 0 ? 1.0f : 0.5f );
            const float w  = k * ( inRenderArea( qx, qy, lo, hi, gh ) ? 1.0f : 0.0f );
            acc            = acc + xyz( Ld4( src, d.GW, d.GH, qx, qy ) ) * w;
            wsum += w;
        }
    }
    const float4 r = f4( acc / wsum, 0.0f );
    St4( dst, d.GW, gx, gy, r );
    if( iter == 0 )
    {
        St4( P.cur[ RG_EXT_PLANE_DIS_GRADIENT_HISTORY ], d.GW, gx, gy, r );
    }
}

// ---------------------------------------------------------- temporal

template< bool F16 >
__global__ __launch_bounds__( THREADS ) void k_temporal( FbPtrs P, Dims d, RgExtFrameParams g )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX;
    const int y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = d.W, H = d.H;
    if( x >= W || y >= H )
    {
        return;
    }
    const v3     unfDiff   = xyz( Ld4( P.cur[ RG_EXT_PLANE_UNFILTERED_DIRECT ], W, H, x, y ) );
    const v3     unfIndir  = xyz( Ld4( P.cur[ RG_EXT_PLANE_UNFILTERED_INDIR ], W, H, x, y ) );
    const v3     unfSpec   = xyz( Ld4( P.cur[ RG_EXT_PLANE_UNFILTERED_SPECULAR ], W, H, x, y ) );
    const float4 rayInfo   = Ld4( P.cur[ RG_EXT_PLANE_VIEW_DIRECTION ], W, H, x, y );
    const v3     rayDir    = xyz( rayInfo );
    const float4 motion    = Ld4( P.cur[ RG_EXT_PLANE_MOTION ], W, H, x, y );
    const float  depth     = Ld1( P.cur[ RG_EXT_PLANE_DEPTH_WORLD ], W, H, x, y );
    const v3     normal    = DecodeNormal( LdU( P.cur[ RG_EXT_PLANE_NORMAL ], W, H, x, y ) );
    const float  roughness = Ld2( P.cur[ RG_EXT_PLANE_METALLIC_ROUGHNESS ], W, H, x, y ).y;
    int          alo, ahi;
    cbRenderArea( x, W, &alo, &ahi );

    const float l = luminance( unfDiff );
    float       px, py;
    PrevScreenPos( motion.x, motion.y, x, y, W, H, &px, &py );
    const int   ppx = int( floorf( px - 0.5f ) ), ppy = int( floorf( py - 0.5f ) );
    const float sx = ( px - 0.5f ) - floorf( px - 0.5f ), sy = ( py - 0.5f ) - floorf( py - 0.5f );
    const float bw[ 2 ][ 2 ] = { { ( 1.0f - sx ) * ( 1.0f - sy ), sx * ( 1.0f - sy ) }, { ( 1.0f - sx ) * sy, sx * sy } };

    float4 hist      = make_float4( 0.f, 0.f, 0.f, 0.f );
    v3     diffPrev  = mk3( 0, 0, 0 ), specPrev = mk3( 0, 0, 0 ), indirPrev = mk3( 0, 0, 0 );
    float2 momPrev   = make_float2( 0.f, 0.f );
    float  weightSum = 0.0f;
#pragma unroll
    for( int yy = 0; yy <= 1; yy++ )
    {
#pragma unroll
        for( int xx = 0; xx <= 1; xx++ )
        {
            const int   qx = ppx + xx, qy = ppy + yy;
            const float dPrev = Ld1( P.prev[ RG_EXT_PLANE_DEPTH_WORLD ], W, H, qx, qy );
            const v3    nPrev = DecodeNormal( LdU( P.prev[ RG_EXT_PLANE_NORMAL ], W, H, qx, qy ) );
            if( inRenderArea( qx, qy, alo, ahi, H ) && TestDepth( depth, dPrev, motion.z ) && TestNormal( normal, nPrev ) )
            {
                const float  w  = bw[ yy ][ xx ];
                const float2 mp = LdP2< F16 >( P.prev[ RG_EXT_PLANE_DIFF_ACCUM_MOMENTS ], W, H, qx, qy );
                const float4 hp = LdP4< F16 >( P.prev[ RG_EXT_PLANE_ACCUM_HISTORY_LENGTH ], W, H, qx, qy );
                diffPrev        = diffPrev + xyz( LdP4< F16 >( P.cur[ RG_EXT_PLANE_DIFF_COLOR_HISTORY ], W, H, qx, qy ) ) * w;
                momPrev.x += w * mp.x;
                momPrev.y += w * mp.y;
                hist.x += w * hp.x;
                hist.y += w * hp.y;
                hist.z += w * hp.z;
                hist.w += w * hp.w;
                specPrev  = specPrev + xyz( LdP4< F16 >( P.prev[ RG_EXT_PLANE_SPEC_ACCUM_COLOR ], W, H, qx, qy ) ) * w;
                indirPrev = indirPrev + xyz( LdP4< F16 >( P.prev[ RG_EXT_PLANE_INDIR_ACCUM ], W, H, qx, qy ) ) * w;
                weightSum += w;
            }
        }
    }

    const float4 grad   = Ld4( P.cur[ RG_EXT_PLANE_DIS_PING_GRADIENT ], d.GW, d.GH, x / STRATA, y / STRATA );
    const float  aDiff  = grad.x;
    const float  aIndir = grad.y;
    const float  aSpec  = grad.z;

    v3    specAccum;
    float specHistAccum, specHitAccum;
    if( weightSum > 0.0001f )
    {
        specPrev                 = specPrev / weightSum;
        const float specHistLen  = hist.z / weightSum;
        const float specHitPrev  = hist.w / weightSum;
        const v3    surfPos      = xyz( Ld4( P.cur[ RG_EXT_PLANE_SURFACE_POSITION ], W, H, x, y ) );
        const v3    surfPosPrev  = xyz( Ld4( P.prev[ RG_EXT_PLANE_SURFACE_POSITION ], W, H, ppx, ppy ) );
        const float nv           = fmaxf( 0.0f, dot( normal, -rayDir ) );
        const float A            = clampf( 1.0f + specHistLen * Pow5( 1.0f - aSpec ), SpecMinAccumFrames( g ), 256.0f );
        const float parallax     = Parallax( g, surfPos, surfPosPrev );
        const float AOnSurf      = SpecAccumSpeed( g, A, roughness, nv, parallax );
        const v3    specOnSurf   = mix3( specPrev, unfSpec, 1.0f / ( 1.0f + AOnSurf ) );
        const float hitOnSurf    = mixf( specHitPrev, rayInfo.w, 1.0f / ( 1.0f + AOnSurf ) );
        specAccum                = specOnSurf;
        specHistAccum            = AOnSurf;
        specHitAccum             = hitOnSurf;
        const float dominant     = SpecularDominantFactor( nv, roughness );
        const v3    virtualPos   = surfPos + rayDir * ( hitOnSurf * dominant );

        float vc[ 4 ], vp[ 4 ], cc[ 4 ], cp[ 4 ];
        MulCM( g.view, virtualPos, 1.0f, vc );
        MulCM( g.viewPrev, virtualPos, 1.0f, vp );
        MulCM( g.projection, mk3( vc[ 0 ], vc[ 1 ], vc[ 2 ] ), vc[ 3 ], cc );
        MulCM( g.projectionPrev, mk3( vp[ 0 ], vp[ 1 ], vp[ 2 ] ), vp[ 3 ], cp );
        const float scx = ( cc[ 0 ] / cc[ 3 ] ) * 0.5f + 0.5f, scy = ( cc[ 1 ] / cc[ 3 ] ) * 0.5f + 0.5f;
        const float spx = ( cp[ 0 ] / cp[ 3 ] ) * 0.5f + 0.5f, spy = ( cp[ 1 ] / cp[ 3 ] ) * 0.5f + 0.5f;
        float       qpx, qpy;
        PrevScreenPos( spx - scx, spy - scy, x, y, W, H, &qpx, &qpy );
        const int   vpx = int( floorf( qpx - 0.5f ) ), vpy = int( floorf( qpy - 0.5f ) );
        const float tx = ( qpx - 0.5f ) - floorf( qpx - 0.5f ), ty = ( qpy - 0.5f ) - floorf( qpy - 0.5f );
        const float bws[ 2 ][ 2 ] = { { ( 1.0f - tx ) * ( 1.0f - ty ), tx * ( 1.0f - ty ) }, { ( 1.0f - tx ) * ty, tx * ty } };
        float       wsv = 0.0f;
#pragma unroll
        for( int yy = 0; yy <= 1; yy++ )
        {
#pragma unroll
            for( int xx = 0; xx <= 1; xx++ )
            {
                const int   qx = vpx + xx, qy = vpy + yy;
                const float dPrev = Ld1( P.prev[ RG_EXT_PLANE_DEPTH_WORLD ], W, H, qx, qy );
                const v3    nPrev = DecodeNormal( LdU( P.prev[ RG_EXT_PLANE_NORMAL ], W, H, qx, qy ) );
                if( inRenderArea( qx, qy, alo, ahi, H ) && TestDepth( depth, dPrev, motion.z ) && TestNormal( normal, nPrev ) )
                {
                    // the reference accumulates the virtual-motion sample but
                    // blends specPrevOnSurf (CmSVGFTemporalAccumulation.comp:262-275)
                    wsv += bws[ yy ][ xx ];
                }
            }
        }
        if( wsv > 0.0001f )
        {
            float       Avirt = SpecAccumSpeed( g, A, roughness, nv, 0.0f );
            const float Amin  = fminf( Avirt, 3.0f * sqrtf( roughness ) );
            float       a     = mixf( 1.0f / ( 1.0f + Amin ), 1.0f / ( 1.0f + Avirt ), 1.0f );
            Avirt             = 1.0f / a - 1.0f;
            const float Ahit  = fminf( Avirt, 16.0f );
            const v3    specVirt = mix3( specPrev, unfSpec, 1.0f / ( 1.0f + Avirt ) );
            a                    = mixf( 1.0f / ( 1.0f + AOnSurf ), 1.0f / ( 1.0f + Avirt ), dominant );
            specAccum            = mix3( specOnSurf, specVirt, dominant );
            specHistAccum        = 1.0f / a - 1.0f;
            specHitAccum         = mixf( specHitPrev, rayInfo.w, 1.0f / ( 1.0f + Ahit ) );
        }
    }
    else
    {
        specAccum     = unfSpec;
        specHistAccum = 1.0f;
        specHitAccum  = rayInfo.w;
    }

    v3     diffAccum, indirAccum;
    float2 diffMom;
    float  diffHist, indirHist;
    if( weightSum > 0.0001f )
    {
        const float inv = 1.0f / weightSum;
        diffPrev        = diffPrev * inv;
        momPrev.x *= inv;
        momPrev.y *= inv;
        indirPrev = indirPrev * inv;
        float dh  = hist.x * inv;
        float ih  = hist.y * inv;

        dh *= Pow10( 1.0f - aDiff );
        dh           = clampf( dh + 1.0f, 1.0f, 256.0f );
        float alphaC = fmaxf( 0.01f, 1.0f / dh );
        float alphaM = fmaxf( 0.01f, 1.0f / dh );
        alphaC       = mixf( alphaC, 1.0f, aDiff );
        alphaM       = mixf( alphaM, 1.0f, aDiff );
        diffAccum    = mix3( diffPrev, unfDiff, alphaC );
        diffMom      = make_float2( mixf( momPrev.x, l, alphaM ), mixf( momPrev.y, l * l, alphaM ) );
        diffHist     = dh;

        ih *= Pow10( 1.0f - aIndir );
        ih         = clampf( ih + 1.0f, 1.0f, 256.0f );
        float aI   = fmaxf( 0.01f, 1.0f / ih );
        aI         = mixf( aI, 1.0f, aIndir );
        indirAccum = mix3( indirPrev, unfIndir, aI );
        indirHist  = ih;
    }
    else
    {
        diffAccum  = unfDiff;
        diffMom    = make_float2( l, l * l );
        diffHist   = 1.0f;
        indirAccum = unfIndir;
        indirHist  = 1.0f;
    }
    St4( P.cur[ RG_EXT_PLANE_DENOISE_GUIDE ], W, x, y, f4( normal, depth ) );
    StP4< F16 >( P.cur[ RG_EXT_PLANE_DIFF_TEMPORARY ], W, x, y, f4( diffAccum, 0.0f ) );
    StP2< F16 >( P.cur[ RG_EXT_PLANE_DIFF_ACCUM_MOMENTS ], W, x, y, diffMom );
    StP4< F16 >( P.cur[ RG_EXT_PLANE_ACCUM_HISTORY_LENGTH ], W, x, y, make_float4( diffHist, indirHist, specHistAccum, specHitAccum ) );
    StP4< F16 >( P.cur[ RG_EXT_PLANE_SPEC_ACCUM_COLOR ], W, x, y, f4( specAccum, 0.0f ) );
    StP4< F16 >( P.cur[ RG_EXT_PLANE_INDIR_ACCUM ], W, x, y, f4( indirAccum, 0.0f ) );
}

// --------------------------------------------------------- anti-firefly

template< bool F16 >
__global__ __launch_bounds__( THREADS ) void k_antifirefly( FbPtrs P, Dims d, uint32_t enabled )
{
    constexpr int R = 2, TW = BX + 2 * R, TH = BY + 2 * R;
    __shared__ v3 s_c[ 3 ][ TH ][ TW ];
    const int     W = d.W, H = d.H;
    const int     bx0 = blockIdx.x * BX - R, by0 = blockIdx.y * BY - R;
    const void*   src[ 3 ] = { P.cur[ RG_EXT_PLANE_DIFF_TEMPORARY ], P.cur[ RG_EXT_PLANE_SPEC_ACCUM_COLOR ], P.cur[ RG_EXT_PLANE_INDIR_ACCUM ] };
    for( int i = threadIdx.x; i < TW * TH; i += THREADS )
    {
        const int tx = i % TW, ty = i / TW;
#pragma unroll
        for( int s = 0; s < 3; s++ )
        {
            s_c[ s ][ ty ][ tx ] = xyz( LdP4< F16 >( src[ s ], W, H, bx0 + tx, by0 + ty ) );
        }
    }
    __syncthreads();
    const int lx = threadIdx.x % BX, ly = threadIdx.x / BX;
    const int x = blockIdx.x * BX + lx, y = blockIdx.y * BY + ly;
    if( x >= W || y >= H )
    {
        return;
    }
    void* dst[ 3 ] = { P.cur[ RG_EXT_PLANE_DIFF_ACCUM_COLOR ], P.cur[ RG_EXT_PLANE_SPEC_PING ], P.cur[ RG_EXT_PLANE_INDIR_PING ] };
#pragma unroll
    for( int s = 0; s < 3; s++ )
    {
        v3 c = s_c[ s ][ ly + R ][ lx + R ];
        if( enabled )
        {
            v3 mn = mk3( 1e6f, 1e6f, 1e6f ), mx = mk3( -1e6f, -1e6f, -1e6f );
            for( int j = -R; j <= R; j++ )
            {
                for( int i = -R; i <= R; i++ )
                {
                    if( i == 0 && j == 0 )
                    {
                        continue;
                    }
                    const v3 o = s_c[ s ][ ly + R + j ][ lx + R + i ];
                    mn         = vmin( mn, o );
                    mx         = vmax( mx, o );
                }
            }
            c = mk3( clampf( c.x, mn.x, mx.x ), clampf( c.y, mn.y, mx.y ), clampf( c.z, mn.z, mx.z ) );
        }
        StP4< F16 >( dst[ s ], W, x, y, f4( c, 0.0f ) );
    }
}

// Writes the denoise guide from NORMAL and DEPTH_WORLD, for the stencil
// passes when they run on their own (the temporal pass writes it in the
// full chain).
__global__ __launch_bounds__( THREADS ) void k_guide( FbPtrs P, Dims d )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX;
    const int y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = d.W, H = d.H;
    if( x >= W || y >= H )
    {
        return;
    }
    St4( P.cur[ RG_EXT_PLANE_DENOISE_GUIDE ], W, x, y,
         f4( DecodeNormal( LdU( P.cur[ RG_EXT_PLANE_NORMAL ], W, H, x, y ) ), Ld1( P.cur[ RG_EXT_PLANE_DEPTH_WORLD ], W, H, x, y ) ) );
}

// ------------------------------------------------------------ variance

template< bool F16 >
__global__ __launch_bounds__( THREADS ) void k_variance( FbPtrs P, Dims d )
{
    constexpr int R = 1, TW = BX + 2 * R, TH = BY + 2 * R;
    __shared__ v3    s_col[ TH ][ TW ];
    __shared__ v3    s_nrm[ TH ][ TW ];
    __shared__ float s_dep[ TH ][ TW ];
    const int        W = d.W, H = d.H;
    const int        bx0 = blockIdx.x * BX - R, by0 = blockIdx.y * BY - R;
    for( int i = threadIdx.x; i < TW * TH; i += THREADS )
    {
        const int tx = i % TW, ty = i / TW;
        s_col[ ty ][ tx ] = xyz( LdP4< F16 >( P.cur[ RG_EXT_PLANE_DIFF_ACCUM_COLOR ], W, H, bx0 + tx, by0 + ty ) );
        LdGuide( P, W, H, bx0 + tx, by0 + ty, s_nrm[ ty ][ tx ], s_dep[ ty ][ tx ] );
    }
    __syncthreads();
    const int lx = threadIdx.x % BX + R, ly = threadIdx.x / BX + R;
    const int x = blockIdx.x * BX + lx - R, y = blockIdx.y * BY + ly - R;
    if( x >= W || y >= H )
    {
        return;
    }
    v3          color = s_col[ ly ][ lx ];
    const float depth = s_dep[ ly ][ lx ];
    void*       out   = P.cur[ RG_EXT_PLANE_DIFF_PING ];
    if( depth < 0.0f || depth > MAX_RAY_LENGTH )
    {
        StP4< F16 >( out, W, x, y, f4( color, 0.0f ) );
        return;
    }
    const float  histLen = LdP4< F16 >( P.cur[ RG_EXT_PLANE_ACCUM_HISTORY_LENGTH ], W, H, x, y ).x;
    const float2 mom     = LdP2< F16 >( P.cur[ RG_EXT_PLANE_DIFF_ACCUM_MOMENTS ], W, H, x, y );
    constexpr float HT   = 4.0f;
    if( histLen > HT )
    {
        float var = fmaxf( 0.0f, mom.y - mom.x * mom.x );
        var /= fmaxf( 1.0f, histLen - HT );
        StP4< F16 >( out, W, x, y, f4( color, var ) );
        return;
    }
    const float l    = luminance( color );
    float       s0   = l + mom.x, s1 = l * l + mom.y;
    float       wsum = 1.0f;
    const v3    n    = s_nrm[ ly ][ lx ];
#pragma unroll
    for( int yy = -1; yy <= 1; yy++ )
    {
#pragma unroll
        for( int xx = -1; xx <= 1; xx++ )
        {
            if( xx == 0 && yy == 0 )
            {
                continue;
            }
            const v3    cq  = s_col[ ly + yy ][ lx + xx ];
            const float lq  = luminance( cq );
            const float dq  = s_dep[ ly + yy ][ lx + xx ];
            const float dm  = fmaxf( depth, dq );
            const float w_z = FastExp( -fabsf( depth - dq ) * ( dm <= 0.0f ? 0.0f : __builtin_amdgcn_rcpf( dm ) ) );
            const float w_n = Pow128( fmaxf( 0.0f, dot( n, s_nrm[ ly + yy ][ lx + xx ] ) ) );
            const float w   = w_z * w_n;
            if( !isnan( w ) )
            {
                s0 += lq * w;
                s1 += lq * lq * w;
                color = color + cq * w;
                wsum += w;
            }
        }
    }
    const float inv = safePositiveRcp( wsum );
    s0 *= inv;
    s1 *= inv;
    color           = color * inv;
    const float var = fmaxf( 0.0f, s1 - s0 * s0 ) * ( 1.0f + 2.0f * ( 1.0f - histLen / HT ) );
    StP4< F16 >( out, W, x, y, f4( color, var ) );
}

// ------------------------------------------------------------- a-trous

struct AtrousPlanes
{
    int srcD, srcS, srcI, dstD, dstS, dstI; // dst = -1: compose
};

// Iteration k's inputs and outputs (CmSVGFAtrous.comp:257-278), continued
// past the fourth iteration when atrousIterations = 5.
AtrousPlanes GetAtrousPlanes( uint32_t k, uint32_t count )
{
    AtrousPlanes a;
    if( k == 0 )
    {
        a = { RG_EXT_PLANE_DIFF_PING, RG_EXT_PLANE_SPEC_PING, RG_EXT_PLANE_INDIR_PING, RG_EXT_PLANE_DIFF_COLOR_HISTORY, RG_EXT_PLANE_SPEC_PONG, RG_EXT_PLANE_INDIR_PONG };
        return a;
    }
    if( k == 1 )
    {
        a.srcD = RG_EXT_PLANE_DIFF_COLOR_HISTORY, a.srcS = RG_EXT_PLANE_SPEC_PONG, a.srcI = RG_EXT_PLANE_INDIR_PONG;
    }
    else if( k % 2 == 0 )
    {
        a.srcD = RG_EXT_PLANE_DIFF_PING, a.srcS = RG_EXT_PLANE_SPEC_PING, a.srcI = RG_EXT_PLANE_INDIR_PING;
    }
    else
    {
        a.srcD = RG_EXT_PLANE_DIFF_PONG, a.srcS = RG_EXT_PLANE_SPEC_PONG, a.srcI = RG_EXT_PLANE_INDIR_PONG;
    }
    if( k + 1 == count )
    {
        a.dstD = a.dstS = a.dstI = -1;
    }
    else if( k % 2 == 1 )
    {
        a.dstD = RG_EXT_PLANE_DIFF_PING, a.dstS = RG_EXT_PLANE_SPEC_PING, a.dstI = RG_EXT_PLANE_INDIR_PING;
    }
    else
    {
        a.dstD = RG_EXT_PLANE_DIFF_PONG, a.dstS = RG_EXT_PLANE_SPEC_PONG, a.dstI = RG_EXT_PLANE_INDIR_PONG;
    }
    return a;
}

__device__ __forceinline__ float RoughnessSquaredToSpecPower( float a )
{
    return fmaxf( 0.01f, 2.0f / ( a * a + 1e-4f ) - 2.0f );
}

// CmSVGFAtrous.comp:284-348: (diffuse + indirect) * rho_d + specular * rho_s,
// times throughput and the material ambient term; regular-layout outputs.
__device__ void Compose( const FbPtrs& P, const Dims& d, int x, int y, v3 diff, v3 spec, v3 indir )
{
    const int W = d.W, H = d.H;
    int       rx, ry;
    cbToRegular( x, y, W, &rx, &ry );
    const v3 albedo     = xyz( Ld4( P.cur[ RG_EXT_PLANE_ALBEDO ], W, H, rx, ry ) );
    const v3 throughput = xyz( Ld4( P.cur[ RG_EXT_PLANE_THROUGHPUT ], W, H, x, y ) );
    v3       ill;
    if( LdU( P.cur[ RG_EXT_PLANE_IS_SKY ], W, H, x, y ) == 0u )
    {
        const float2 mr = Ld2( P.cur[ RG_EXT_PLANE_METALLIC_ROUGHNESS ], W, H, x, y );
        float        t  = clampf( ( mr.y - 0.5f ) / 0.25f, 0.0f, 1.0f );
        t               = t * t * ( 3.0f - 2.0f * t );
        const v3 specular = mix3( spec, indir, t );
        const v3 ros      = mix3( mk3( 0.04f, 0.04f, 0.04f ), albedo, mr.x );
        const v3 rod      = albedo * ( 1.0f - mr.x );
        ill               = ( diff + indir ) * rod + specular * ros;
        ill               = ill * throughput;
        const float la    = luminance( albedo );
        const float amb   = la > 0.02f ? 1.0f : 1.0f - square( ( la - 0.02f ) / 0.02f );
        ill               = ill * amb;
    }
    else
    {
        ill = albedo * throughput;
    }
    St1( P.cur[ RG_EXT_PLANE_HISTOGRAM_INPUT ], W, rx, ry, fmaxf( 0.0f, luminance( ill ) ) );
    St4( P.cur[ RG_EXT_PLANE_PRE_FINAL ], W, rx, ry, make_float4( fmaxf( 0.0f, ill.x ), fmaxf( 0.0f, ill.y ), fmaxf( 0.0f, ill.z ), 0.0f ) );
}

// One edge-avoiding a-trous step; tap data comes from `Tap`.
template< bool F16, typename TapFn >
__device__ void AtrousPixel( const FbPtrs& P, const Dims& d, int x, int y, int k, int step, const AtrousPlanes& ap,
                             v3 c_n, float c_depth, float c_grad, float c_rough, float4 c_diff, v3 c_spec, v3 c_indir,
                             float pvar, TapFn&& tap )
{
    const int W = d.W, H = d.H;
    int       alo, ahi;
    cbRenderArea( x, W, &alo, &ahi );
    v3    outD = xyz( c_diff );
    float outV = c_diff.w;
    v3    outS = c_spec, outI = c_indir;

    const float l      = luminance( outD );
    const float wLum   = 1.0f / ( 4.0f * pvar + 0.00001f );
    const float wRough = clampf( c_rough * 30.0f - float( k ), 0.0f, 1.0f );
    const float hSpec  = LdP4< F16 >( P.cur[ RG_EXT_PLANE_ACCUM_HISTORY_LENGTH ], W, H, x, y ).z;
    float       nws    = clampf( RoughnessSquaredToSpecPower( c_rough * c_rough ), 8.0f, 1024.0f );
    nws *= clampf( hSpec / 8.0f, 0.0f, 1.0f );

    // 1 / max(grad * (|dx| + |dy|), 0.01) of the edge and diagonal taps
    const float zInv1 = 1.0f / fmaxf( c_grad, 0.01f ), zInv2 = 1.0f / fmaxf( c_grad * 2.0f, 0.01f );
    float ws = 1.0f, wss = 1.0f, wsi = 1.0f;
#pragma unroll 1
    for( int yy = -1; yy <= 1; yy++ )
    {
#pragma unroll
        for( int xx = -1; xx <= 1; xx++ )
        {
            if( xx == 0 && yy == 0 )
            {
                continue;
            }
            const int qx = x + xx * step, qy = y + yy * step;
            v3        nq, sq, iq;
            float4    cq;
            float     dq, rq;
            tap( xx, yy, qx, qy, nq, dq, rq, cq, sq, iq );
            const float lq  = luminance( xyz( cq ) );
            const float nn  = fmaxf( 0.0f, dot( c_n, nq ) );
            const float w_z = fabsf( c_depth - dq ) * ( xx != 0 && yy != 0 ? zInv2 : zInv1 );
            const float w_n = Pow128( nn );
            const float w_l = fabsf( l - lq ) * wLum;
            float       w_r = fmaxf( 0.0f, 1.0f - 10.0f * fabsf( c_rough - rq ) ) * wRough;
            if( nws > 0.0f )
            {
                w_r *= FastPow( nn, nws );
            }
            const float kw     = ( xx == 0 ? 1.0f : 0.5f ) * ( yy == 0 ? 1.0f : 0.5f );
            const float inside = inRenderArea( qx, qy, alo, ahi, H ) ? 1.0f : 0.0f;
            const float wBase  = FastExp( -w_z * w_z ) * w_n * kw * inside;
            const float wD = wBase * FastExp( -w_l ), wS = wBase * w_r, wI = wBase;
            outD = outD + xyz( cq ) * wD;
            outS = outS + sq * wS;
            outI = outI + iq * wI;
            outV += cq.w * wD * wD;
            ws += wD;
            wss += wS;
            wsi += wI;
        }
    }
    const float iw = 1.0f / ws;
    outD           = outD * iw;
    outV *= iw * iw;
    outS = outS * ( 1.0f / wss );
    outI = outI * ( 1.0f / wsi );
    if( ap.dstD >= 0 )
    {
        StP4< F16 >( P.cur[ ap.dstD ], W, x, y, f4( outD, outV ) );
        StP4< F16 >( P.cur[ ap.dstS ], W, x, y, f4( outS, 0.0f ) );
        StP4< F16 >( P.cur[ ap.dstI ], W, x, y, f4( outI, 0.0f ) );
    }
    else
    {
        Compose( P, d, x, y, outD, outS, outI );
    }
}

// Iteration 0: the 3x3 footprint (plus the 3x3 Gaussian of the variance) is
// staged in LDS (CmSVGFAtrous_Iter0.comp:96-163).
template< bool F16 >
__global__ __launch_bounds__( THREADS ) void k_atrous0( FbPtrs P, Dims d, AtrousPlanes ap )
{
    constexpr int R = 1, TW = BX + 2 * R, TH = BY + 2 * R;
    __shared__ float4 s_diff[ TH ][ TW ];
    __shared__ v3     s_spec[ TH ][ TW ];
    __shared__ v3     s_ind[ TH ][ TW ];
    __shared__ v3     s_nrm[ TH ][ TW ];
    __shared__ float  s_dep[ TH ][ TW ];
    __shared__ float  s_rgh[ TH ][ TW ];
    const int         W = d.W, H = d.H;
    const int         bx0 = blockIdx.x * BX - R, by0 = blockIdx.y * BY - R;
    for( int i = threadIdx.x; i < TW * TH; i += THREADS )
    {
        const int tx = i % TW, ty = i / TW, gx = bx0 + tx, gy = by0 + ty;
        s_diff[ ty ][ tx ] = LdP4< F16 >( P.cur[ ap.srcD ], W, H, gx, gy );
        s_spec[ ty ][ tx ] = xyz( LdP4< F16 >( P.cur[ ap.srcS ], W, H, gx, gy ) );
        s_ind[ ty ][ tx ]  = xyz( LdP4< F16 >( P.cur[ ap.srcI ], W, H, gx, gy ) );
        LdGuide( P, W, H, gx, gy, s_nrm[ ty ][ tx ], s_dep[ ty ][ tx ] );
        s_rgh[ ty ][ tx ]  = Ld2( P.cur[ RG_EXT_PLANE_METALLIC_ROUGHNESS ], W, H, gx, gy ).y;
    }
    __syncthreads();
    const int lx = threadIdx.x % BX + R, ly = threadIdx.x / BX + R;
    const int x = lx - R + blockIdx.x * BX, y = ly - R + blockIdx.y * BY;
    if( x >= W || y >= H )
    {
        return;
    }
    const float depth = s_dep[ ly ][ lx ];
    if( depth < 0.0f || depth > MAX_RAY_LENGTH )
    {
        const float4 z = make_float4( 0.f, 0.f, 0.f, 0.f );
        if( ap.dstD >= 0 )
        {
            StP4< F16 >( P.cur[ ap.dstD ], W, x, y, z );
            StP4< F16 >( P.cur[ ap.dstS ], W, x, y, z );
            StP4< F16 >( P.cur[ ap.dstI ], W, x, y, z );
        }
        else
        {
            Compose( P, d, x, y, mk3( 0, 0, 0 ), mk3( 0, 0, 0 ), mk3( 0, 0, 0 ) );
        }
        return;
    }
    float r = 0.0f;
#pragma unroll
    for( int yy = -1; yy <= 1; yy++ )
    {
#pragma unroll
        for( int xx = -1; xx <= 1; xx++ )
        {
            r += s_diff[ ly + yy ][ lx + xx ].w * ( ( xx == 0 ? 0.5f : 0.25f ) * ( yy == 0 ? 0.5f : 0.25f ) );
        }
    }
    const float pvar = sqrtf( fmaxf( r, 0.0f ) );
    StP1< F16 >( P.cur[ RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE ], W, x, y, pvar );
    AtrousPixel< F16 >( P, d, x, y, 0, 1, ap, s_nrm[ ly ][ lx ], depth, Ld1( P.cur[ RG_EXT_PLANE_DEPTH_GRAD ], W, H, x, y ),
                 s_rgh[ ly ][ lx ], s_diff[ ly ][ lx ], s_spec[ ly ][ lx ], s_ind[ ly ][ lx ], pvar,
                 [ & ]( int xx, int yy, int, int, v3& nq, float& dq, float& rq, float4& cq, v3& sq, v3& iq ) {
                     nq = s_nrm[ ly + yy ][ lx + xx ];
                     dq = s_dep[ ly + yy ][ lx + xx ];
                     rq = s_rgh[ ly + yy ][ lx + xx ];
                     cq = s_diff[ ly + yy ][ lx + xx ];
                     sq = s_spec[ ly + yy ][ lx + xx ];
                     iq = s_ind[ ly + yy ][ lx + xx ];
                 } );
}

// Iterations >= 1 (steps 2, 4, 8, 16): taps gathered from global memory.
// Each pixel is computed on its own, so the workgroup tile (TBX x
// THREADS / TBX pixels) changes only which tap rows share a cache.
template< bool F16, int TBX >
__global__ __launch_bounds__( THREADS ) void k_atrousN( FbPtrs P, Dims d, AtrousPlanes ap, int k )
{
    const int x = blockIdx.x * TBX + threadIdx.x % TBX;
    const int y = blockIdx.y * ( THREADS / TBX ) + threadIdx.x / TBX;
    const int W = d.W, H = d.H;
    if( x >= W || y >= H )
    {
        return;
    }
    const float4 cd    = LdP4< F16 >( P.cur[ ap.srcD ], W, H, x, y );
    const v3     cs    = xyz( LdP4< F16 >( P.cur[ ap.srcS ], W, H, x, y ) );
    const v3     ci    = xyz( LdP4< F16 >( P.cur[ ap.srcI ], W, H, x, y ) );
    v3           n;
    float        depth;
    LdGuide( P, W, H, x, y, n, depth );
    if( depth < 0.0f || depth > MAX_RAY_LENGTH )
    {
        if( ap.dstD >= 0 )
        {
            StP4< F16 >( P.cur[ ap.dstD ], W, x, y, cd );
            StP4< F16 >( P.cur[ ap.dstS ], W, x, y, f4( cs, 0.0f ) );
            StP4< F16 >( P.cur[ ap.dstI ], W, x, y, f4( ci, 0.0f ) );
        }
        else
        {
            Compose( P, d, x, y, xyz( cd ), cs, ci );
        }
        return;
    }
    const float rough = Ld2( P.cur[ RG_EXT_PLANE_METALLIC_ROUGHNESS ], W, H, x, y ).y;
    const float pvar  = LdP1< F16 >( P.cur[ RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE ], W, H, x, y );
    AtrousPixel< F16 >( P, d, x, y, k, 1 << k, ap, n, depth, Ld1( P.cur[ RG_EXT_PLANE_DEPTH_GRAD ], W, H, x, y ), rough, cd, cs, ci, pvar,
                 [ & ]( int, int, int qx, int qy, v3& nq, float& dq, float& rq, float4& cq, v3& sq, v3& iq ) {
                     LdGuide( P, W, H, qx, qy, nq, dq );
                     rq = Ld2( P.cur[ RG_EXT_PLANE_METALLIC_ROUGHNESS ], W, H, qx, qy ).y;
                     cq = LdP4< F16 >( P.cur[ ap.srcD ], W, H, qx, qy );
                     sq = xyz( LdP4< F16 >( P.cur[ ap.srcS ], W, H, qx, qy ) );
                     iq = xyz( LdP4< F16 >( P.cur[ ap.srcI ], W, H, qx, qy ) );
                 } );
}

Dims GetDims( const FramebufferSet& fb )
{
    return Dims{ int( fb.width ), int( fb.height ), int( fb.gw ), int( fb.gh ) };
}

void CheckParams( const FramebufferSet& fb, const RgExtFrameParams& p )
{
    RG_REQUIRE( p.renderWidth == fb.width && p.renderHeight == fb.height, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                "frame params render size differs from the framebuffer size" );
    RG_REQUIRE( p.atrousIterations == 4 || p.atrousIterations == 5, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "atrousIterations must be 4 or 5" );
}

} // namespace

void DenoiseGradientAtrous( FramebufferSet& fb, const RgExtFrameParams& p, uint32_t iter, hipStream_t s )
{
    CheckParams( fb, p );
    RG_REQUIRE( iter < 4, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "gradient a-trous iteration must be < 4" );
    const Dims d = GetDims( fb );
    k_grad_atrous<<< dim3( DivUp( d.GW, BX ), DivUp( d.GH, BY ) ), THREADS, 0, s >>>( fb.Ptrs(), d, int( iter ) );
    CheckLaunch( "k_grad_atrous" );
}

void DenoiseTemporal( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s )
{
    CheckParams( fb, p );
    const Dims d = GetDims( fb );
    const dim3 grid( DivUp( d.W, BX ), DivUp( d.H, BY ) );
    if( fb.Half() )
    {
        k_temporal< true ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, p );
    }
    else
    {
        k_temporal< false ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, p );
    }
    CheckLaunch( "k_temporal" );
}

void DenoiseAntiFirefly( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s )
{
    CheckParams( fb, p );
    const Dims d = GetDims( fb );
    const dim3 grid( DivUp( d.W, BX ), DivUp( d.H, BY ) );
    if( fb.Half() )
    {
        k_antifirefly< true ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, p.antiFireflyEnabled );
    }
    else
    {
        k_antifirefly< false ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, p.antiFireflyEnabled );
    }
    CheckLaunch( "k_antifirefly" );
}

namespace
{
void WriteGuide( FramebufferSet& fb, hipStream_t s )
{
    const Dims d = GetDims( fb );
    k_guide<<< dim3( DivUp( d.W, BX ), DivUp( d.H, BY ) ), THREADS, 0, s >>>( fb.Ptrs(), d );
    CheckLaunch( "k_guide" );
}
} // namespace

void DenoiseVariance( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s, bool writeGuide )
{
    CheckParams( fb, p );
    const Dims d = GetDims( fb );
    const dim3 grid( DivUp( d.W, BX ), DivUp( d.H, BY ) );
    if( writeGuide )
    {
        WriteGuide( fb, s );
    }
    if( fb.Half() )
    {
        k_variance< true ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d );
    }
    else
    {
        k_variance< false ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d );
    }
    CheckLaunch( "k_variance" );
}

namespace
{
template< int TBX >
void LaunchAtrousNTile( FramebufferSet& fb, const Dims& d, const AtrousPlanes& ap, uint32_t k, hipStream_t s )
{
    const dim3 grid( DivUp( d.W, TBX ), DivUp( d.H, THREADS / TBX ) );
    if( fb.Half() )
    {
        k_atrousN< true, TBX ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, ap, int( k ) );
    }
    else
    {
        k_atrousN< false, TBX ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, ap, int( k ) );
    }
    CheckLaunch( "k_atrousN" );
}

void LaunchAtrousN( FramebufferSet& fb, const Dims& d, const AtrousPlanes& ap, uint32_t k, uint32_t tileW, hipStream_t s )
{
    switch( tileW )
    {
        case 16: LaunchAtrousNTile< 16 >( fb, d, ap, k, s ); break;
        case 64: LaunchAtrousNTile< 64 >( fb, d, ap, k, s ); break;
        case 128: LaunchAtrousNTile< 128 >( fb, d, ap, k, s ); break;
        default: LaunchAtrousNTile< 32 >( fb, d, ap, k, s ); break;
    }
}
} // namespace

void DenoiseAtrous( FramebufferSet& fb, const RgExtFrameParams& p, uint32_t k, hipStream_t s, bool writeGuide )
{
    CheckParams( fb, p );
    RG_REQUIRE( k < p.atrousIterations, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a-trous iteration out of range" );
    if( writeGuide )
    {
        WriteGuide( fb, s );
    }
    const Dims         d  = GetDims( fb );
    const AtrousPlanes ap = GetAtrousPlanes( k, p.atrousIterations );
    const dim3         grid( DivUp( d.W, BX ), DivUp( d.H, BY ) );
    if( k == 0 )
    {
        if( fb.Half() )
        {
            k_atrous0< true ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, ap );
        }
        else
        {
            k_atrous0< false ><<< grid, THREADS, 0, s >>>( fb.Ptrs(), d, ap );
        }
        CheckLaunch( "k_atrous0" );
    }
    else
    {
        LaunchAtrousN( fb, d, ap, k, fb.atrousTileW, s );
    }
}

void SetAtrousTile( FramebufferSet& fb, uint32_t tileWidth )
{
    bool known = false;
    for( uint32_t w : ATROUS_TILE_WIDTHS )
    {
        known = known || w == tileWidth;
    }
    RG_REQUIRE( known, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "a-trous tile width must be 16, 32, 64 or 128" );
    fb.atrousTileW = tileWidth;
}

uint32_t AutotuneAtrous( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s, float* msOut )
{
    CheckParams( fb, p );
    WriteGuide( fb, s );
    const Dims         d  = GetDims( fb );
    const AtrousPlanes ap = GetAtrousPlanes( 1, p.atrousIterations );
    hipEvent_t         e0, e1;
    RG_HIP_CHECK( hipEventCreate( &e0 ) );
    RG_HIP_CHECK( hipEventCreate( &e1 ) );
    uint32_t best   = fb.atrousTileW;
    float    bestMs = 1e30f;
    try
    {
        for( uint32_t i = 0; i < 4; i++ )
        {
            const uint32_t w = ATROUS_TILE_WIDTHS[ i ];
            LaunchAtrousN( fb, d, ap, 1, w, s ); // warm-up
            float ms = 1e30f;
            for( int r = 0; r < 3; r++ )
            {
                RG_HIP_CHECK( hipEventRecord( e0, s ) );
                LaunchAtrousN( fb, d, ap, 1, w, s );
                RG_HIP_CHECK( hipEventRecord( e1, s ) );
                RG_HIP_CHECK( hipEventSynchronize( e1 ) );
                float t = 0.0f;
                RG_HIP_CHECK( hipEventElapsedTime( &t, e0, e1 ) );
                ms = std::min( ms, t );
            }
            if( msOut )
            {
                msOut[ i ] = ms;
            }
            if( ms < bestMs )
            {
                bestMs = ms;
                best   = w;
            }
        }
    }
    catch( ... )
    {
        ( void )hipEventDestroy( e0 );
        ( void )hipEventDestroy( e1 );
        throw;
    }
    RG_HIP_CHECK( hipEventDestroy( e0 ) );
    RG_HIP_CHECK( hipEventDestroy( e1 ) );
    fb.atrousTileW = best;
    return best;
}

void Denoise( FramebufferSet& fb, const RgExtFrameParams& p, hipStream_t s )
{
    for( uint32_t i = 0; i < 4; i++ )
    {
        DenoiseGradientAtrous( fb, p, i, s );
    }
    DenoiseTemporal( fb, p, s );
    DenoiseAntiFirefly( fb, p, s );
    // the temporal pass has written the guide
    DenoiseVariance( fb, p, s, false );
    for( uint32_t k = 0; k < p.atrousIterations; k++ )
    {
        DenoiseAtrous( fb, p, k, s, false );
    }
}

} // namespace rg
