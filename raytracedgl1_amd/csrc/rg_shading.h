// Device-side lights, BRDF, sampling and reservoirs of the path tracer.
//
// Restates Source/Shaders/Light.h:225-313 (light sampling), BRDF.h:27-221
// (Lambert, GGX, VNDF sampling), Random.h:34-200 (disk/hemisphere warps,
// orthonormal basis) and Reservoir.h:24-146 (RIS reservoirs).
#pragma once

#include "rg_device_math.h"
#include "rg_scene.h"

namespace rg
{

constexpr float RG_PI          = 3.14159265358979323846f;
constexpr float MAX_RAY_LEN    = 10000.0f; // MAX_RAY_LENGTH
constexpr float MIN_GGX_ROUGH  = 0.005f;   // MIN_GGX_ROUGHNESS (BRDF.h:110)

// Random.h:24-32
constexpr uint32_t SALT_DIFF_BOUNCE( uint32_t i ) { return 8 + i; }
constexpr uint32_t SALT_SPEC_BOUNCE( uint32_t i ) { return 12 + i; }
constexpr uint32_t SALT_LIGHT_POINT           = 20;
constexpr uint32_t SALT_LIGHT_GRID_BASE       = 24;
constexpr uint32_t SALT_INITIAL_RESERVOIRS    = 48;
constexpr uint32_t SALT_LIGHT_CHOOSE_DIRECT   = 72;
constexpr uint32_t SALT_LIGHT_CHOOSE_INDIRECT = 96;
constexpr uint32_t SALT_RESAMPLE_INDIRECT_BASE = 132;

struct Surface
{
    v3       position;
    uint32_t instCustomIndex;
    v3       normal;
    float    roughness;
    v3       albedo;
    bool     isSky;
    v3       specularColor;
    v3       toViewerDir;
};

struct LightSample
{
    v3    position;
    v3    color;
    float dw;
};

struct Reservoir
{
    uint32_t selected;
    float    targetPdf;
    float    weightSum;
    uint32_t M;
};

RG_HD v3 safeNormalize( v3 v )
{
    const float len = length( v );
    return len > 0.001f ? v / len : mk3( 0.0f, 1.0f, 0.0f );
}

// Hardware reciprocal, square root and reciprocal square root (v_rcp_f32,
// v_sqrt_f32, v_rsq_f32: 1 ulp) for the light sampling and BRDF evaluated
// per RIS candidate: a correctly rounded divide or square root is about ten
// instructions, and the candidate loop runs up to 32 times per pixel.  Zero,
// infinite and NaN inputs give the same specials as the operations they
// replace.
__device__ __forceinline__ float fastRcp( float x )
{
    return __builtin_amdgcn_rcpf( x );
}
__device__ __forceinline__ float fastSqrt( float x )
{
    return __builtin_amdgcn_sqrtf( x );
}
__device__ __forceinline__ float fastLength( v3 a )
{
    return fastSqrt( dot( a, a ) );
}
__device__ __forceinline__ v3 fastNormalize( v3 a )
{
    return a * __builtin_amdgcn_rsqf( dot( a, a ) );
}
__device__ __forceinline__ v3 fastSafeNormalize( v3 v )
{
    const float len = fastLength( v );
    return len > 0.001f ? v * fastRcp( len ) : mk3( 0.0f, 1.0f, 0.0f );
}
__device__ __forceinline__ float fastSafePositiveRcp( float f )
{
    return f <= 0.0f ? 0.0f : fastRcp( f );
}

// ----------------------------------------------------------- sampling

// Random.h:36-51
__device__ __forceinline__ void sampleDisk( float radius, float u1, float u2, float* x, float* y )
{
    u1 *= 0.99f;
    u2 *= 0.99f;
    const float r   = radius * sqrtf( u1 );
    const float phi = 2.0f * RG_PI * u2;
    *x              = r * cosf( phi );
    *y              = r * sinf( phi );
}

// "Building an Orthonormal Basis, Revisited" (Random.h:116-134)
__device__ __forceinline__ void revisedONB( v3 n, v3* b1, v3* b2 )
{
    if( n.z < 0.0f )
    {
        const float a = 1.0f / ( 1.0f - n.z );
        const float b = n.x * n.y * a;
        *b1           = mk3( 1.0f - n.x * n.x * a, -b, n.x );
        *b2           = mk3( b, n.y * n.y * a - 1.0f, -n.y );
    }
    else
    {
        const float a = 1.0f / ( 1.0f + n.z );
        const float b = -n.x * n.y * a;
        *b1           = mk3( 1.0f - n.x * n.x * a, b, -n.x );
        *b2           = mk3( b, 1.0f - n.y * n.y * a, -n.y );
    }
}

// Random.h:168-200 ("Ray Tracing Gems", ch. 16)
__device__ __forceinline__ v3 sampleOrientedHemisphere( v3 n, float u1, float u2, float* oneOverPdf )
{
    float       a   = 1.0f - 2.0f * u1;
    float       b   = sqrtf( 1.0f - a * a );
    const float phi = 2.0f * RG_PI * u2;
    a *= 0.98f;
    b *= 0.98f;
    v3 r        = normalize( mk3( n.x + b * cosf( phi ), n.y + b * sinf( phi ), n.z + a ) );
    *oneOverPdf = RG_PI * safePositiveRcp( dot( r, n ) );
    return r;
}

// The light-point random numbers of a pixel and the terms of sampleDisk and
// sampleOrientedHemisphere that depend only on them.  Every candidate light
// of a pixel is sampled at the same (u, v) (Light.h:225-313 takes one
// pointRnd per surface), so the sines and cosines are taken once per pixel;
// each term is the same expression the warps evaluate.
struct LightPoint
{
    float diskR, diskCos, diskSin;      // sampleDisk( 1, pu, pv ): sqrt(0.99 u1), cos / sin(2 pi 0.99 u2)
    float hemiX, hemiY, hemiZ;          // sampleOrientedHemisphere( n, pu, pv ) = normalize( n + hemi )
    float triU1, triU2;                 // the triangle light's 0.99-scaled pair
};

__device__ __forceinline__ LightPoint MakeLightPoint( float pu, float pv )
{
    LightPoint lp;
    {
        const float u1 = pu * 0.99f, u2 = pv * 0.99f;
        const float phi = 2.0f * RG_PI * u2;
        lp.diskR        = sqrtf( u1 );
        lp.diskCos      = cosf( phi );
        lp.diskSin      = sinf( phi );
        lp.triU1        = u1;
        lp.triU2        = u2;
    }
    {
        float       a   = 1.0f - 2.0f * pu;
        float       b   = sqrtf( 1.0f - a * a );
        const float phi = 2.0f * RG_PI * pv;
        a *= 0.98f;
        b *= 0.98f;
        lp.hemiX = b * cosf( phi );
        lp.hemiY = b * sinf( phi );
        lp.hemiZ = a;
    }
    return lp;
}

// ---------------------------------------------------------------- BRDF

// BRDF.h:83-103; loose = FORCE_EVALBRDF_GGX_LOOSE (alpha >= 0.2), defined
// for the indirect-illumination shaders only (RtRaygenIndirect.inl:26).
__device__ __forceinline__ float D_GGX( float nm, float alpha, bool loose )
{
    if( loose )
    {
        alpha = fmaxf( 0.2f, alpha );
    }
    const float a2 = alpha * alpha;
    nm             = fmaxf( 0.0f, nm );
    return a2 * ( 1.0f / RG_PI ) * fastRcp( square( nm * nm * ( a2 - 1.0f ) + 1.0f ) );
}

__device__ __forceinline__ float G1_GGX( float ns, float alpha )
{
    return 2.0f * ns * safePositiveRcp( ns * ( 2.0f - alpha ) + alpha );
}

// BRDF.h:116-144
__device__ __forceinline__ v3 evalBRDFSmithGGX( v3 n, v3 v, v3 l, float alpha, v3 specColor, bool loose )
{
    alpha          = fmaxf( alpha, MIN_GGX_ROUGH );
    const float nl = fmaxf( dot( n, l ), 0.0f );
    if( nl <= 0.0f )
    {
        return mk3( 0, 0, 0 );
    }
    const float f  = powf( 1.0f - fmaxf( nl, 0.0f ), 5.0f );
    const v3    F  = specColor + ( mk3( 1, 1, 1 ) - specColor ) * f;
    const v3    h  = fastNormalize( v + l );
    const float D  = D_GGX( dot( n, h ), alpha, loose );
    const float nv = fmaxf( dot( n, v ), 0.0f );
    const float G2 = 0.5f * fastRcp( mixf( 2.0f * nl * nv, nl + nv, alpha ) );
    return F * ( G2 * D );
}

// Heitz, "Sampling the GGX Distribution of Visible Normals" (BRDF.h:153-196)
__device__ __forceinline__ v3 sampleGGXVNDF( v3 v, float alpha, float u1, float u2, float* oneOverPdf, bool loose )
{
    alpha = fmaxf( alpha, MIN_GGX_ROUGH );
    u1 *= 0.98f;
    u2 *= 0.98f;
    const v3    Vh    = normalize( mk3( alpha * v.x, alpha * v.y, v.z ) );
    const float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    const v3    T1    = lensq > 0.0f ? mk3( -Vh.y, Vh.x, 0.0f ) * ( 1.0f / sqrtf( lensq ) ) : mk3( 1, 0, 0 );
    const v3    T2    = cross( Vh, T1 );
    const float r     = sqrtf( u1 );
    const float phi   = 2.0f * RG_PI * u2;
    const float t1    = r * cosf( phi );
    float       t2    = r * sinf( phi );
    const float s     = 0.5f * ( 1.0f + Vh.z );
    t2                = ( 1.0f - s ) * sqrtf( 1.0f - t1 * t1 ) + s * t2;
    const v3 Nh       = T1 * t1 + T2 * t2 + Vh * sqrtf( fmaxf( 0.0f, 1.0f - t1 * t1 - t2 * t2 ) );
    const v3 Ne       = normalize( mk3( alpha * Nh.x, alpha * Nh.y, fmaxf( 0.02f, Nh.z ) ) );
    const float D     = D_GGX( Ne.z, alpha, loose );
    const float G1    = G1_GGX( v.z, alpha );
    *oneOverPdf       = v.z * safePositiveRcp( G1 * fmaxf( 0.0f, dot( v, Ne ) ) * D );
    return Ne;
}

// BRDF.h:205-221: world-space reflection direction and its 1/pdf.
__device__ __forceinline__ v3 sampleSmithGGX( v3 n, v3 v, float alpha, float u1, float u2, float* oneOverPdf, bool loose )
{
    v3 b0, b1;
    revisedONB( n, &b0, &b1 );
    const v3 ve = mk3( dot( b0, v ), dot( b1, v ), dot( n, v ) );
    const v3 m  = sampleGGXVNDF( ve, alpha, u1, u2, oneOverPdf, loose );
    // reflect(-ve, m) = -ve - 2 dot(m, -ve) m
    const v3 l = -ve + m * ( 2.0f * dot( m, ve ) );
    *oneOverPdf *= 4.0f * dot( ve, m );
    return b0 * l.x + b1 * l.y + n * l.z;
}

// ---------------------------------------------------------------- lights

// Light.h:127-154
__device__ __forceinline__ float safeSolidAngle( float a )
{
    return ( a > 0.0f && !isnan( a ) && !isinf( a ) ) ? clampf( a, 0.0f, 4.0f * RG_PI ) : 0.0f;
}
__device__ __forceinline__ float solidAngleForSphere( float radius, float dist )
{
    const float sinT = radius >= dist ? 1.0f : radius * fastRcp( dist );
    const float cosT = fastSqrt( 1.0f - sinT * sinT );
    return safeSolidAngle( 2.0f * RG_PI * ( 1.0f - cosT ) );
}
__device__ __forceinline__ float lightColorWeight( v3 c )
{
    return clampf( luminance( c ) * 0.1f + 0.9f, 1.0f, 10.0f );
}

__device__ __forceinline__ v3 L3( const float* a )
{
    return mk3( a[ 0 ], a[ 1 ], a[ 2 ] );
}

// getLightWeight (Light.h:162-204, :286-296), for the light grid
__device__ float lightWeight( const LightEncoded& e, v3 cellCenter, float cellRadius, float polyFactor )
{
    const v3 color = L3( e.color );
    switch( e.lightType )
    {
        case LIGHT_TYPE_DIRECTIONAL: return lightColorWeight( color );
        case LIGHT_TYPE_SPHERE:
            return lightColorWeight( color ) * solidAngleForSphere( e.data0[ 3 ], fmaxf( length( L3( e.data0 ) - cellCenter ), cellRadius ) );
        case LIGHT_TYPE_TRIANGLE: {
            const v3    p0 = L3( e.data0 ), p1 = L3( e.data1 ), p2 = L3( e.data2 );
            const v3    c  = p0 / 3.0f + p1 / 3.0f + p2 / 3.0f;
            const float r  = length( p0 - c ) / 3.0f + length( p1 - c ) / 3.0f + length( p2 - c ) / 3.0f;
            v3          nrm = mk3( e.data0[ 3 ], e.data1[ 3 ], e.data2[ 3 ] );
            nrm             = nrm / length( nrm );
            const float front = dot( nrm, cellCenter - c ) > -cellRadius ? 1.0f : 0.0f;
            return lightColorWeight( color ) * solidAngleForSphere( r, fmaxf( length( c - cellCenter ), cellRadius ) ) * front;
        }
        case LIGHT_TYPE_SPOT: {
            const v3    c     = L3( e.data0 );
            const float front = dot( L3( e.data1 ), cellCenter - c ) > -cellRadius ? 1.0f : 0.0f;
            return lightColorWeight( color ) * solidAngleForSphere( e.data0[ 3 ], fmaxf( length( c - cellCenter ), cellRadius ) ) * front;
        }
        default: return 0.0f;
    }
}

// sampleLight (Light.h:225-313), at the pixel's light point.
__device__ LightSample sampleLight( const LightEncoded& e, v3 surfPos, const LightPoint& lp, float polyFactor )
{
    LightSample r;
    r.position = mk3( 0, 0, 0 );
    r.color    = mk3( 0, 0, 0 );
    r.dw       = 0.0f;
    switch( e.lightType )
    {
        case LIGHT_TYPE_DIRECTIONAL: {
            const v3    dir  = L3( e.data0 );
            const float rad  = sinf( fmaxf( 0.01f, e.data0[ 3 ] ) );
            const float rr   = rad * lp.diskR;
            v3          b0, b1;
            revisedONB( dir, &b0, &b1 );
            const v3 ln = fastNormalize( dir + b0 * ( rr * lp.diskCos ) + b1 * ( rr * lp.diskSin ) );
            r.position  = surfPos - ln * MAX_RAY_LEN;
            r.color     = L3( e.color );
            r.dw        = 1.0f;
            return r;
        }
        case LIGHT_TYPE_SPHERE: {
            // the hemisphere around the direction to the surface
            const v3    c   = L3( e.data0 );
            const v3    d   = c - surfPos;
            const float len = fastLength( d );
            const v3    ln  = fastNormalize( d * -fastRcp( len ) + mk3( lp.hemiX, lp.hemiY, lp.hemiZ ) );
            r.position      = c + ln * e.data0[ 3 ];
            r.color         = L3( e.color );
            r.dw            = solidAngleForSphere( e.data0[ 3 ], len );
            return r;
        }
        case LIGHT_TYPE_TRIANGLE: {
            const v3 p0 = L3( e.data0 ), p1 = L3( e.data1 ), p2 = L3( e.data2 );
            v3       nrm = mk3( e.data0[ 3 ], e.data1[ 3 ], e.data2[ 3 ] );
            const float nl = fastLength( nrm );
            nrm            = nrm * fastRcp( nl );
            const float area = nl * 0.5f;
            const float beta = 1.0f - sqrtf( lp.triU1 ), gamma = ( 1.0f - beta ) * lp.triU2, alpha = 1.0f - beta - gamma;
            r.position       = p0 * alpha + p1 * beta + p2 * gamma;
            v3          tl   = surfPos - r.position;
            const float dist = fastLength( tl );
            tl               = tl * fastRcp( dist );
            const float ll   = fmaxf( dot( nrm, tl ), 0.0f );
            r.color          = L3( e.color ) * powf( ll, polyFactor );
            r.dw             = safeSolidAngle( area * ( fabsf( dot( nrm, tl ) ) * fastRcp( square( dist ) ) ) );
            return r;
        }
        case LIGHT_TYPE_SPOT: {
            const v3    c = L3( e.data0 ), dir = L3( e.data1 );
            const float rr = e.data0[ 3 ] * lp.diskR;
            v3          b0, b1;
            revisedONB( dir, &b0, &b1 );
            r.position      = c + b0 * ( rr * lp.diskCos ) + b1 * ( rr * lp.diskSin );
            const v3    d   = c - surfPos;
            const float len = fastLength( d );
            const float cosA = fmaxf( -dot( dir, d ) * fastRcp( len ), 0.0f );
            float       t    = clampf( ( cosA - e.data2[ 1 ] ) * fastRcp( e.data2[ 0 ] - e.data2[ 1 ] ), 0.0f, 1.0f );
            t                = t * t * ( 3.0f - 2.0f * t );
            r.color          = L3( e.color ) * square( t );
            r.dw             = solidAngleForSphere( e.data0[ 3 ], len );
            return r;
        }
        default: return r;
    }
}

// shade (RaygenCommon.h:364-380)
__device__ __forceinline__ void shadeLight( const Surface& s, const LightSample& lt, float oneOverPdf, bool loose, v3* diff, v3* spec )
{
    const v3    l  = fastSafeNormalize( lt.position - s.position );
    const float nl = dot( s.normal, l );
    if( nl <= 0.0f )
    {
        *diff = *spec = mk3( 0, 0, 0 );
        return;
    }
    *diff = lt.color * ( lt.dw * nl * ( 1.0f / RG_PI ) );
    *spec = ( lt.color * ( lt.dw * nl ) ) * evalBRDFSmithGGX( s.normal, s.toViewerDir, l, s.roughness, s.specularColor, loose );
    *diff = *diff * oneOverPdf;
    *spec = *spec * oneOverPdf;
}

__device__ __forceinline__ float targetPdf( const Surface& s, const LightSample& lt, bool loose )
{
    v3 d, sp;
    shadeLight( s, lt, 1.0f, loose, &d, &sp );
    return luminance( d + sp );
}

// ------------------------------------------------------------ reservoirs

__device__ __forceinline__ Reservoir emptyReservoir()
{
    return Reservoir{ LIGHT_INDEX_NONE, 0.0f, 0.0f, 0u };
}
__device__ __forceinline__ bool isValid( const Reservoir& r )
{
    return r.selected != LIGHT_INDEX_NONE;
}
__device__ __forceinline__ float selectedSampleWeight( const Reservoir& r )
{
    return safePositiveRcp( r.targetPdf ) * ( r.weightSum / float( r.M > 1u ? r.M : 1u ) );
}
__device__ __forceinline__ void normalizeReservoir( Reservoir& r, uint32_t maxM )
{
    r.weightSum /= float( r.M > 1u ? r.M : 1u );
    r.M = r.M < maxM ? r.M : maxM;
    r.weightSum *= float( r.M );
}
__device__ __forceinline__ void updateReservoir( Reservoir& r, uint32_t light, float tp, float oneOverSourcePdf, float rnd )
{
    const float w = tp * oneOverSourcePdf;
    r.weightSum += w;
    r.M += 1;
    if( rnd * r.weightSum < w )
    {
        r.selected  = light;
        r.targetPdf = tp;
    }
}
__device__ __forceinline__ void updateCombined( Reservoir& c, const Reservoir& b, float rnd )
{
    const float w = b.weightSum;
    c.weightSum += w;
    c.M += b.M;
    if( rnd * c.weightSum < w )
    {
        c.selected  = b.selected;
        c.targetPdf = b.targetPdf;
    }
}
__device__ __forceinline__ void updateCombinedNewSurf( Reservoir& c, const Reservoir& b, float tpB, float rnd )
{
    const float w = tpB * safePositiveRcp( b.targetPdf ) * b.weightSum;
    c.weightSum += w;
    c.M += b.M;
    if( rnd * c.weightSum < w )
    {
        c.selected  = b.selected;
        c.targetPdf = tpB;
    }
}
// packReservoir / unpackReservoir (Reservoir.h:116-146)
__device__ __forceinline__ uint4 packReservoir( const Reservoir& r )
{
    if( !isinf( r.weightSum ) && !isnan( r.weightSum ) )
    {
        return make_uint4( ( ( r.M < 65535u ? r.M : 65535u ) << 16 ) | ( r.selected < 65535u ? r.selected : 65535u ),
                           __float_as_uint( r.targetPdf ), __float_as_uint( r.weightSum ), 0u );
    }
    return make_uint4( LIGHT_INDEX_NONE, 0u, 0u, 0u );
}
__device__ __forceinline__ Reservoir unpackReservoir( uint4 p )
{
    return Reservoir{ p.x & 65535u, __uint_as_float( p.y ), __uint_as_float( p.z ), ( p.x >> 16 ) & 65535u };
}

} // namespace rg
