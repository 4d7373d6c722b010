// Ray-traced passes of a frame, as wavefronts around the traversal kernel.
//
//   primary      RtRaygenPrimary.rgen -> RaygenPrimary.inl:280-348 (G-buffer)
//   light grid   CmLightGridBuild.comp:46-105
//   initial      RtInitialReservoirs.rgen -> calcInitialReservoir
//                (RaygenCommon.h:396-486), directional-light visibility ray
//   direct       RtRaygenDirect.rgen -> selectLight_Direct / traceDirect-
//                Illumination (RaygenCommon.h:503-674), shadow ray
//   indirect     RtRaygenIndirect.inl:66-240 (GGX-VNDF bounce, shadowed
//                direct light at the hit, second diffuse bounce) and the
//                shading of RtRaygenIndirect.inl:402-433
//
// Each pass writes rays into a buffer, the persistent traversal kernel
// resolves them, and a shading kernel consumes the hits; per-pixel state
// between the two halves of a pass lives in "pending" planes.  Pixels that
// need no ray emit an empty ray (tmax < tmin) that the traversal rejects at
// the root.
#include "rg_denoise.h"
#include "rg_pathtrace.h"
#include "rg_shading.h"
#include "rg_textures.h"

#include <mutex>

namespace rg
{
namespace
{

constexpr int BX = 32, BY = 8, THREADS = 256;
constexpr int GRID_X = 16, GRID_Y = 16, GRID_Z = 16, GRID_CELL = 128; // LIGHT_GRID_SIZE_*, LIGHT_GRID_CELL_SIZE
constexpr uint32_t GRID_ENTRIES = GRID_X * GRID_Y * GRID_Z * GRID_CELL;

struct SceneView
{
    const float*             tris;
    const float*             normals;
    const uint32_t*          triGeom;
    const GeomRecord*        geoms;
    const RgPrimitiveVertex* verts;
    const uint32_t*          indices;
    TexView                  tex;
};

// How a hit samples its albedo texture: the three texture LOD sources of
// HitInfo.inl (primary rays: gradients from ray differentials; reflection /
// refraction rays: ray cones; bounces: getBounceLOD).
enum TexMode : uint32_t
{
    TQ_LOD    = 0, // a = lod
    TQ_GRAD   = 1, // dbx, dby: barycentric deltas of the neighbouring pixels' rays
    TQ_CONE   = 2, // a = cone width, dir = ray direction
    TQ_BOUNCE = 3, // a = origin roughness, b = mip bias
};
struct TexQuery
{
    uint32_t mode = TQ_LOD;
    float    a = 0.0f, b = 0.0f;
    v3       dbx, dby, dir;
};

struct LightView
{
    const LightEncoded* lights;
    uint32_t            arraySize;
    const uint32_t*     prevToCur;
    uint32_t            prevCount;
    const uint4*        grid;
    const uint4*        gridPrev;
    const uint32_t*     blueNoise;
};

__device__ __forceinline__ LightEncoded GetLight( const LightView& lv, uint32_t i )
{
    if( i < lv.arraySize )
    {
        return lv.lights[ i ];
    }
    LightEncoded none{};
    return none;
}

// ------------------------------------------------------------ helpers

template< typename T >
__device__ __forceinline__ T* Plane( const FbPtrs& P, int plane, bool prev = false )
{
    return reinterpret_cast< T* >( prev ? P.prev[ plane ] : P.cur[ plane ] );
}

__device__ __forceinline__ v3 DecodeNormal( uint32_t p )
{
    float sp, cp, st, ct;
    sincospif( unitQuotient< 65535u >( p >> 16 ), &sp, &cp );
    sincospif( float( p & 0xFFFFu ) / 32768.0f, &st, &ct );
    return mk3( sp * ct, sp * st, cp );
}

// encodeNormal (Utils.h:108-124)
__device__ __forceinline__ uint32_t EncodeNormal( v3 n )
{
    const float phi   = acosf( clampf( n.z, -1.0f, 1.0f ) );
    float       theta = atan2f( n.y, n.x );
    theta             = theta < 0.0f ? theta + 2.0f * RG_PI : theta;
    const uint32_t j  = uint32_t( roundf( phi * 65535.0f / RG_PI ) );
    const uint32_t k  = uint32_t( roundf( theta * 65536.0f / ( 2.0f * RG_PI ) ) ) % 65536u;
    return ( j << 16 ) | k;
}

// encodeE5B9G9R9 / decodeE5B9G9R9 (Utils.h:150-207, EXT_texture_shared_exponent):
// 9-bit mantissas, shared 5-bit exponent with bias 15; floor(log2) and the
// powers of two are taken from the exponent bits, so both are exact.
__device__ __forceinline__ uint32_t EncodeE5B9G9R9( v3 c )
{
    const float r = fminf( fmaxf( c.x, 0.0f ), 65408.0f );
    const float g = fminf( fmaxf( c.y, 0.0f ), 65408.0f );
    const float b = fminf( fmaxf( c.z, 0.0f ), 65408.0f );
    const float m = fmaxf( r, fmaxf( g, b ) );
    if( !( m > 0.0f ) )
    {
        return 0u;
    }
    const int   e    = int( ( __float_as_uint( m ) >> 23 ) & 255u ) - 127;
    const int   expP = ( e > -16 ? e : -16 ) + 16;
    const int   maxS = int( rintf( m * ldexpf( 1.0f, 24 - expP ) ) );
    const int   ex   = maxS != 512 ? expP : expP + 1;
    const float sc   = ldexpf( 1.0f, 24 - ex );
    return ( uint32_t( ex ) << 27 ) | ( uint32_t( rintf( b * sc ) ) << 18 ) | ( uint32_t( rintf( g * sc ) ) << 9 ) | uint32_t( rintf( r * sc ) );
}

__device__ __forceinline__ v3 DecodeE5B9G9R9( uint32_t p )
{
    const float sc = ldexpf( 1.0f, int( p >> 27 ) - 24 );
    return mk3( sc * float( p & 511u ), sc * float( ( p >> 9 ) & 511u ), sc * float( ( p >> 18 ) & 511u ) );
}

__device__ __forceinline__ v3 MulM4( const float* m, v3 v, float w, float* wOut )
{
    const v3 r = mk3( m[ 0 ] * v.x + m[ 4 ] * v.y + m[ 8 ] * v.z + m[ 12 ] * w,
                      m[ 1 ] * v.x + m[ 5 ] * v.y + m[ 9 ] * v.z + m[ 13 ] * w,
                      m[ 2 ] * v.x + m[ 6 ] * v.y + m[ 10 ] * v.z + m[ 14 ] * w );
    *wOut      = m[ 3 ] * v.x + m[ 7 ] * v.y + m[ 11 ] * v.z + m[ 15 ] * w;
    return r;
}

// getRayDir (ShaderCommonGLSLFunc.h:460-468)
__device__ __forceinline__ v3 GetRayDir( const RgExtFrameParams& p, float u, float v )
{
    u = u * 2.0f - 1.0f;
    v = v * 2.0f - 1.0f;
    float    tw;
    const v3 t     = MulM4( p.invProjection, mk3( u, v, 1.0f ), 1.0f, &tw );
    const v3 local = fabsf( tw ) < 0.001f ? t : t / tw;
    float    w;
    return MulM4( p.invView, normalize( local ), 0.0f, &w );
}

__device__ __forceinline__ v3 UnpackColor( uint32_t c )
{
    return mk3( unitQuotient< 255u >( c & 0xFFu ), unitQuotient< 255u >( ( c >> 8 ) & 0xFFu ), unitQuotient< 255u >( ( c >> 16 ) & 0xFFu ) );
}

// getSky / adjustSky (RaygenCommon.h:272-306) for RG_SKY_TYPE_COLOR
__device__ __forceinline__ v3 SkyColor( const RgExtFrameParams& p )
{
    const v3    c = mk3( p.skyColorDefault[ 0 ], p.skyColorDefault[ 1 ], p.skyColorDefault[ 2 ] );
    const float g = luminance( c );
    return mix3( mk3( g, g, g ), c, p.skyColorSaturation ) * p.skyColorMultiplier;
}

__device__ __forceinline__ v3 SpecularColor( v3 albedo, float metallic )
{
    return mix3( mk3( 0.04f, 0.04f, 0.04f ), albedo, metallic );
}

__device__ __forceinline__ v3 Xform34( const float* m, v3 p )
{
    return mk3( m[ 0 ] * p.x + m[ 1 ] * p.y + m[ 2 ] * p.z + m[ 3 ],
                m[ 4 ] * p.x + m[ 5 ] * p.y + m[ 6 ] * p.z + m[ 7 ],
                m[ 8 ] * p.x + m[ 9 ] * p.y + m[ 10 ] * p.z + m[ 11 ] );
}

__device__ __forceinline__ RgExtRay MakeRay( v3 o, v3 d, float tmin, float tmax )
{
    RgExtRay r;
    r.origin[ 0 ] = o.x, r.origin[ 1 ] = o.y, r.origin[ 2 ] = o.z, r.tmin = tmin;
    r.direction[ 0 ] = d.x, r.direction[ 1 ] = d.y, r.direction[ 2 ] = d.z, r.tmax = tmax;
    return r;
}
__device__ __forceinline__ RgExtRay NullRay()
{
    return MakeRay( mk3( 0, 0, 0 ), mk3( 0, 1, 0 ), 1.0f, -1.0f );
}

struct HitInfo
{
    v3       position;
    v3       normal;
    v3       albedo;
    v3       emission;
    float    roughness;
    float    metallic;
    uint32_t geom;
    uint32_t localPrim;
};

// getBounceLOD (HitInfo.inl:160-175), "Ray Traced Reflections in
// 'Wolfenstein: Youngblood'"
__device__ float BounceLod( float roughness, float viewDist, float hitDist, float screenWidth, float bias )
{
    const float k     = ( 1.0f - roughness ) * 0.9f + 0.1f;
    const float range = 300.0f * ( k * k * k * k );
    const float fx    = sqrtf( clampf( viewDist / range, 0.0f, 1.0f ) );
    const float fy    = sqrtf( clampf( hitDist / range, 0.0f, 1.0f ) );
    return fmaxf( log2f( 3840.0f / screenWidth ), 0.0f ) + fx * 10.0f + fy * 10.0f + bias;
}

// Layer-0 albedo with its texture (processAlbedo*, HitInfo.inl:28-128, for
// one opaque layer): colour * texture, saturated.
// Texture coordinates of a hit and the LOD source of its query: one
// sampler for all maps of a material (HitInfo.inl:28-128, :386-447).
struct MaterialCoords
{
    float2 t[ 3 ];
    float  u, v;                     // interpolated texture coordinates
    float  dux, dvx, duy, dvy;       // gradients (TQ_GRAD, TQ_CONE)
    float  lod;                      // TQ_LOD, TQ_BOUNCE
};

__device__ __forceinline__ MaterialCoords GetMaterialCoords( const SceneView& sc, const RgExtFrameParams& p, const GeomRecord& gr, uint32_t localPrim,
                                                             float u, float v, v3 p0, v3 p1, v3 p2, v3 position, v3 normal, v3 rayOrigin,
                                                             const TexQuery& q )
{
    MaterialCoords m;
#pragma unroll
    for( int k = 0; k < 3; k++ )
    {
        const uint32_t idx = gr.baseIndex == 0xFFFFFFFFu ? 3 * localPrim + k : sc.indices[ gr.baseIndex + 3 * localPrim + k ];
        const RgPrimitiveVertex& vx = sc.verts[ gr.baseVertex + idx ];
        m.t[ k ] = make_float2( vx.texCoord[ 0 ], vx.texCoord[ 1 ] );
    }
    const float b0 = 1.0f - u - v;
    m.u            = m.t[ 0 ].x * b0 + m.t[ 1 ].x * u + m.t[ 2 ].x * v;
    m.v            = m.t[ 0 ].y * b0 + m.t[ 1 ].y * u + m.t[ 2 ].y * v;
    m.dux = m.dvx = m.duy = m.dvy = 0.0f;
    m.lod = 0.0f;
    if( q.mode == TQ_GRAD )
    {
        m.dux = m.t[ 0 ].x * q.dbx.x + m.t[ 1 ].x * q.dbx.y + m.t[ 2 ].x * q.dbx.z;
        m.dvx = m.t[ 0 ].y * q.dbx.x + m.t[ 1 ].y * q.dbx.y + m.t[ 2 ].y * q.dbx.z;
        m.duy = m.t[ 0 ].x * q.dby.x + m.t[ 1 ].x * q.dby.y + m.t[ 2 ].x * q.dby.z;
        m.dvy = m.t[ 0 ].y * q.dby.x + m.t[ 1 ].y * q.dby.y + m.t[ 2 ].y * q.dby.z;
    }
    else if( q.mode == TQ_CONE )
    {
        // getTriangleUVDerivativesFromRayCone (RayCone.h:86-124)
        const float quadArea = length( cross( p1 - p0, p2 - p0 ) );
        const float pcw      = q.a / fabsf( dot( q.dir, normal ) );
        const float ratio    = pcw * pcw / quadArea;
        const float uvArea   = fabsf( ( m.t[ 1 ].x - m.t[ 0 ].x ) * ( m.t[ 2 ].y - m.t[ 0 ].y ) - ( m.t[ 2 ].x - m.t[ 0 ].x ) * ( m.t[ 1 ].y - m.t[ 0 ].y ) );
        m.dux = m.dvy = sqrtf( uvArea * ratio );
    }
    else if( q.mode == TQ_BOUNCE )
    {
        const v3 cam = mk3( p.cameraPosition[ 0 ], p.cameraPosition[ 1 ], p.cameraPosition[ 2 ] );
        m.lod        = BounceLod( q.a, length( position - cam ), length( position - rayOrigin ), float( p.renderWidth ), q.b );
    }
    else
    {
        m.lod = q.a;
    }
    return m;
}

// One map at the query's LOD; gradients scaled by `gradScale` (the normal
// map's suppressDetails).
__device__ __forceinline__ float4 SampleMaterial( const SceneView& sc, uint32_t id, const MaterialCoords& m, const TexQuery& q, float gradScale )
{
    if( q.mode == TQ_GRAD || q.mode == TQ_CONE )
    {
        return SampleGrad( sc.tex, id, m.u, m.v, m.dux * gradScale, m.dvx * gradScale, m.duy * gradScale, m.dvy * gradScale );
    }
    return SampleLod( sc.tex, id, m.u, m.v, m.lod );
}

// getTangent (VertexData.inl:249-269) from the model-space vertices, then
// into world space by the model matrix's 3x3 part (:488-494): xyz tangent,
// w handedness.
__device__ float4 TriangleTangent( const SceneView& sc, const GeomRecord& gr, uint32_t localPrim, const MaterialCoords& m )
{
    v3 pos[ 3 ], nsum = mk3( 0.f, 0.f, 0.f );
#pragma unroll
    for( int k = 0; k < 3; k++ )
    {
        const uint32_t idx = gr.baseIndex == 0xFFFFFFFFu ? 3 * localPrim + k : sc.indices[ gr.baseIndex + 3 * localPrim + k ];
        const RgPrimitiveVertex& vx = sc.verts[ gr.baseVertex + idx ];
        pos[ k ] = mk3( vx.position[ 0 ], vx.position[ 1 ], vx.position[ 2 ] );
        nsum     = nsum + mk3( vx.normal[ 0 ], vx.normal[ 1 ], vx.normal[ 2 ] );
    }
    const v3     e1 = pos[ 1 ] - pos[ 0 ], e2 = pos[ 2 ] - pos[ 0 ];
    const float2 u1 = make_float2( m.t[ 1 ].x - m.t[ 0 ].x, m.t[ 1 ].y - m.t[ 0 ].y );
    const float2 u2 = make_float2( m.t[ 2 ].x - m.t[ 0 ].x, m.t[ 2 ].y - m.t[ 0 ].y );
    const float  invDet = 1.0f / ( u1.x * u2.y - u2.x * u1.y );
    const v3     tg = normalize( ( e1 * u2.y - e2 * u1.y ) * invDet );
    const v3     bt = normalize( ( e2 * u1.x - e1 * u2.x ) * invDet );
    const float  w  = dot( cross( safeNormalize( nsum ), tg ), bt ) > 0.0f ? 1.0f : -1.0f;
    const float* M  = gr.model;
    return make_float4( M[ 0 ] * tg.x + M[ 1 ] * tg.y + M[ 2 ] * tg.z, M[ 4 ] * tg.x + M[ 5 ] * tg.y + M[ 6 ] * tg.z,
                        M[ 8 ] * tg.x + M[ 9 ] * tg.y + M[ 10 ] * tg.z, w );
}

// Material maps of a textured hit (HitInfo.inl:28-128, :386-447): albedo =
// colour x texture; roughness / metallic from the ORM map's g / b; emission
// from the emission map; the normal map (primary and reflected hits only)
// bends the shading normal by normalMapStrength.
__device__ void ApplyMaterial( const SceneView& sc, const RgExtFrameParams& p, const GeomRecord& gr, uint32_t localPrim, float u, float v, v3 p0,
                               v3 p1, v3 p2, v3 rayOrigin, const TexQuery& q, HitInfo& h, float* rough, float* metal, bool* emissiveMap )
{
    const MaterialCoords m = GetMaterialCoords( sc, p, gr, localPrim, u, v, p0, p1, p2, h.position, h.normal, rayOrigin, q );
    if( TexValid( sc.tex, gr.texture ) )
    {
        const float4 c   = SampleMaterial( sc, gr.texture, m, q, 1.0f );
        const v3     col = UnpackColor( gr.color );
        h.albedo         = mk3( saturate( col.x * c.x ), saturate( col.y * c.y ), saturate( col.z * c.z ) );
    }
    if( TexValid( sc.tex, gr.texOrm ) )
    {
        const float4 orm = SampleMaterial( sc, gr.texOrm, m, q, 1.0f );
        *rough           = orm.y;
        *metal           = orm.z;
    }
    if( TexValid( sc.tex, gr.texEmissive ) )
    {
        const float4 e = SampleMaterial( sc, gr.texEmissive, m, q, 1.0f );
        h.emission     = mk3( e.x, e.y, e.z );
        *emissiveMap   = true;
    }
    if( ( q.mode == TQ_GRAD || q.mode == TQ_CONE ) && TexValid( sc.tex, gr.texNormal ) )
    {
        // less detail in normal maps for better denoising: gradients x 5
        const float4 nm  = SampleMaterial( sc, gr.texNormal, m, q, q.mode == TQ_GRAD ? 5.0f : 1.0f );
        const float  nx  = nm.x * 2.0f - 1.0f, ny = nm.y * 2.0f - 1.0f;
        const float4 tg  = TriangleTangent( sc, gr, localPrim, m );
        const v3     T   = mk3( tg.x, tg.y, tg.z );
        const v3     B   = cross( h.normal, T ) * tg.w;
        const v3     nn  = safeNormalize( T * nx + B * ny + h.normal );
        h.normal         = safeNormalize( mix3( h.normal, nn, p.normalMapStrength ) );
    }
}

// getHitInfo* (HitInfo.inl:204-466): albedo is the primitive colour (layer
// 0) times its albedo texture if it has one, normals interpolated or exact.
template< bool TEX = false >
__device__ HitInfo GetHit( const SceneView& sc, const RgExtFrameParams& p, uint32_t prim, float u, float v, v3 rayOrigin,
                           const TexQuery& q = TexQuery{} )
{
    HitInfo           h;
    const uint32_t    g  = sc.triGeom[ prim ];
    const GeomRecord& gr = sc.geoms[ g ];
    const float*      t  = sc.tris + size_t( prim ) * 9;
    const float*      n  = sc.normals + size_t( prim ) * 9;
    const float       b0 = 1.0f - u - v;
    const v3          p0 = mk3( t[ 0 ], t[ 1 ], t[ 2 ] ), p1 = mk3( t[ 3 ], t[ 4 ], t[ 5 ] ), p2 = mk3( t[ 6 ], t[ 7 ], t[ 8 ] );
    h.position           = p0 * b0 + p1 * u + p2 * v;
    if( ( gr.flags & RG_MESH_PRIMITIVE_FORCE_EXACT_NORMALS ) == 0 )
    {
        h.normal = normalize( mk3( n[ 0 ], n[ 1 ], n[ 2 ] ) * b0 + mk3( n[ 3 ], n[ 4 ], n[ 5 ] ) * u + mk3( n[ 6 ], n[ 7 ], n[ 8 ] ) * v );
    }
    else
    {
        h.normal = safeNormalize( cross( p1 - p0, p2 - p0 ) );
        if( dot( h.normal, h.position - rayOrigin ) > 0.0f )
        {
            h.normal = -h.normal;
        }
    }
    const v3 c  = UnpackColor( gr.color );
    h.albedo    = mk3( saturate( c.x ), saturate( c.y ), saturate( c.z ) );
    float rough = gr.roughness, metal = gr.metallic;
    bool  emissiveMap = false;
    if( TEX && ( gr.texture != TEX_NONE || gr.texOrm != TEX_NONE || gr.texNormal != TEX_NONE || gr.texEmissive != TEX_NONE ) )
    {
        ApplyMaterial( sc, p, gr, prim - gr.firstTri, u, v, p0, p1, p2, rayOrigin, q, h, &rough, &metal, &emissiveMap );
    }
    h.roughness = fmaxf( rough, fmaxf( p.minRoughness, MIN_GGX_ROUGH ) );
    h.metallic  = metal;
    if( !emissiveMap )
    {
        h.emission = h.albedo * gr.emissive;
    }
    h.geom      = g;
    h.localPrim = prim - gr.firstTri;
    return h;
}

// fetchGbufferSurface (Surface.inl:44-69)
__device__ Surface FetchSurface( const FbPtrs& P, int W, int x, int y )
{
    Surface        s;
    const size_t   i = size_t( y ) * W + x;
    s.isSky          = Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ i ] != 0u;
    if( s.isSky )
    {
        return s;
    }
    int rx, ry;
    cbToRegular( x, y, W, &rx, &ry );
    const float4 alb   = Plane< float4 >( P, RG_EXT_PLANE_ALBEDO )[ size_t( ry ) * W + rx ];
    const float4 pos   = Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION )[ i ];
    const float2 mr    = Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ i ];
    const float4 vd    = Plane< float4 >( P, RG_EXT_PLANE_VIEW_DIRECTION )[ i ];
    s.albedo           = mk3( alb.x, alb.y, alb.z );
    s.position         = mk3( pos.x, pos.y, pos.z );
    s.instCustomIndex  = __float_as_uint( pos.w );
    s.specularColor    = SpecularColor( s.albedo, mr.x );
    s.roughness        = mr.y;
    s.normal           = DecodeNormal( Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ i ] );
    s.toViewerDir      = mk3( -vd.x, -vd.y, -vd.z );
    return s;
}

// hitInfoToSurface_Indirect (Surface.inl:96-108)
__device__ Surface HitToSurface( const HitInfo& h, v3 rayDir, uint32_t custom )
{
    Surface s;
    s.position        = h.position;
    s.instCustomIndex = custom;
    s.normal          = h.normal;
    s.roughness       = h.roughness;
    s.albedo          = h.albedo;
    s.isSky           = false;
    s.specularColor   = SpecularColor( h.albedo, h.metallic );
    s.toViewerDir     = -rayDir;
    return s;
}

// rndBlueNoise8 (Random.h:229-238) over the seeded table
__device__ __forceinline__ void BlueNoise8( const uint32_t* table, uint32_t seed, uint32_t salt, float out[ 4 ] )
{
    const uint32_t layer = ( ( seed >> 14 ) + salt ) % 128u;
    const uint32_t oy    = ( seed >> 7 ) & 127u;
    const uint32_t ox    = seed & 127u;
    const uint32_t t     = table[ ( layer * 128u + oy ) * 128u + ox ];
    for( int k = 0; k < 4; k++ )
    {
        out[ k ] = float( ( t >> ( 8 * k ) ) & 0xFFu ) / 255.0f;
    }
}

// --------------------------------------------------------- light grid

struct Grid
{
    v3    minExt, maxExt, delta;
    float cellRadius;
};

// LightGrid.h:31-71
__device__ __forceinline__ Grid MakeGrid( const RgExtFrameParams& p )
{
    Grid        g;
    const float d = p.cellWorldSize;
    g.delta       = mk3( d, d, d );
    const v3 whole = mk3( d * GRID_X, d * GRID_Y, d * GRID_Z );
    const v3 center = mk3( p.cameraPosition[ 0 ], p.cameraPosition[ 1 ], p.cameraPosition[ 2 ] ) + g.delta * 0.5f;
    g.minExt       = center - whole * 0.5f;
    g.maxExt       = center + whole * 0.5f;
    g.cellRadius   = length( g.delta ) * 0.5f;
    return g;
}
__device__ __forceinline__ bool InsideGrid( const Grid& g, v3 p )
{
    return p.x > g.minExt.x && p.y > g.minExt.y && p.z > g.minExt.z && p.x < g.maxExt.x && p.y < g.maxExt.y && p.z < g.maxExt.z;
}
__device__ __forceinline__ int CellArrayIndex( const Grid& g, v3 wp )
{
    int c[ 3 ];
    const v3 r = ( wp - g.minExt );
    c[ 0 ]     = min( max( int( r.x / g.delta.x ), 0 ), GRID_X );
    c[ 1 ]     = min( max( int( r.y / g.delta.y ), 0 ), GRID_Y );
    c[ 2 ]     = min( max( int( r.z / g.delta.z ), 0 ), GRID_Z );
    return GRID_CELL * ( c[ 0 ] + c[ 1 ] * GRID_X + c[ 2 ] * GRID_X * GRID_Y );
}

__global__ __launch_bounds__( THREADS ) void k_light_grid( RgExtFrameParams p, LightView lv, uint4* gridOut )
{
    const int arrayIndex = blockIdx.x * THREADS + threadIdx.x;
    if( arrayIndex >= int( GRID_ENTRIES ) )
    {
        return;
    }
    const Grid g = MakeGrid( p );
    const int  c = arrayIndex / GRID_CELL;
    const int  ci[ 3 ] = { c % GRID_X, ( c % ( GRID_X * GRID_Y ) ) / GRID_X, c / ( GRID_X * GRID_Y ) };
    const v3   cellCenter = g.minExt + g.delta * mk3( ci[ 0 ] + 0.5f, ci[ 1 ] + 0.5f, ci[ 2 ] + 0.5f );
    const uint32_t seed   = randomSeed( arrayIndex, 0, p.frameId );
    uint32_t       salt   = SALT_LIGHT_GRID_BASE;

    Reservoir regular = emptyReservoir();
    for( int i = 0; i < 8; i++ )
    {
        const float    rnd = rnd16( seed, salt++ );
        const uint32_t xi  = 1u + min( uint32_t( rnd * float( p.lightCount ) ), p.lightCount - 1u );
        const float    tp  = lightWeight( GetLight( lv, xi ), cellCenter, g.cellRadius, p.polyLightSpotlightFactor );
        updateReservoir( regular, xi, tp, float( p.lightCount ), rnd16( seed, salt++ ) );
    }
    // temporal reuse of the previous frame's cell sample (LIGHT_GRID_TEMPORAL)
    Reservoir temporal = unpackReservoir( lv.gridPrev[ arrayIndex ] );
    temporal.M         = 1;
    normalizeReservoir( temporal, regular.M * 20 );
    float tpTemporal = 0.0f;
    if( temporal.selected != LIGHT_INDEX_NONE )
    {
        const uint32_t cur = temporal.selected < lv.prevCount ? lv.prevToCur[ temporal.selected ] : 0xFFFFFFFFu;
        if( cur != 0xFFFFFFFFu && cur != LIGHT_INDEX_NONE )
        {
            tpTemporal        = lightWeight( GetLight( lv, cur ), cellCenter, g.cellRadius, p.polyLightSpotlightFactor );
            temporal.selected = cur;
        }
    }
    const float rndRis   = rnd16( seed, salt++ );
    Reservoir   combined = regular;
    updateCombinedNewSurf( combined, temporal, tpTemporal, rndRis );
    normalizeReservoir( combined, 1 );
    // ShLightInCell: {selected, targetPdf, weightSum} (M is 1 by construction)
    gridOut[ arrayIndex ] = make_uint4( combined.selected, __float_as_uint( combined.targetPdf ), __float_as_uint( combined.weightSum ), 0u );
}

// calcInitialReservoir (RaygenCommon.h:396-486) up to the directional-light
// visibility test.  Returns the regular and directional reservoirs and the
// salt to continue with.
__device__ void InitialCandidates( const RgExtFrameParams& p, const LightView& lv, const Surface& surf, uint32_t seed, uint32_t& salt,
                                   const LightPoint& lp, bool loose, Reservoir& regular, Reservoir& dir )
{
    const Grid g = MakeGrid( p );
    regular      = emptyReservoir();
    if( InsideGrid( g, surf.position ) )
    {
        float r4[ 4 ];
        rnd8_4( seed, salt++, r4 );
        v3 jp = surf.position + ( mk3( r4[ 0 ], r4[ 1 ], r4[ 2 ] ) * 2.0f - mk3( 1, 1, 1 ) ) * g.cellRadius;
        jp    = vmin( vmax( jp, g.minExt ), g.maxExt );
        const int base = CellArrayIndex( g, jp );
        for( uint32_t i = 0; i < p.restirInitialCandidates; i++ )
        {
            const float rnd = rnd16( seed, salt++ );
            const int   idx = base + min( max( int( rnd * GRID_CELL ), 0 ), GRID_CELL - 1 );
            // the index can reach one cell past the grid (worldToCell clamps to
            // the size, LightGrid.h:80-86): such a slot reads as empty
            const uint4 cell = idx < int( GRID_ENTRIES ) ? lv.grid[ idx ] : make_uint4( LIGHT_INDEX_NONE, 0u, 0u, 0u );
            Reservoir   r    = Reservoir{ cell.x, __uint_as_float( cell.y ), __uint_as_float( cell.z ), 1u };
            if( !isValid( r ) )
            {
                continue;
            }
            const float       oneOverSource = r.weightSum * safePositiveRcp( r.targetPdf );
            const LightSample ls            = sampleLight( GetLight( lv, r.selected ), surf.position, lp, p.polyLightSpotlightFactor );
            const float       tp            = targetPdf( surf, ls, loose );
            updateReservoir( regular, r.selected, tp, oneOverSource, rnd16( seed, salt++ ) );
        }
    }
    else
    {
        for( uint32_t i = 0; i < p.restirInitialCandidates; i++ )
        {
            const float       rnd = rnd16( seed, salt++ );
            const uint32_t    xi  = 1u + min( uint32_t( rnd * float( p.lightCount ) ), p.lightCount - 1u );
            const LightSample ls  = sampleLight( GetLight( lv, xi ), surf.position, lp, p.polyLightSpotlightFactor );
            const float       tp  = targetPdf( surf, ls, loose );
            updateReservoir( regular, xi, tp, float( p.lightCount ), rnd16( seed, salt++ ) );
        }
    }
    normalizeReservoir( regular, 1 );
    dir = emptyReservoir();
    if( p.directionalLightExists )
    {
        const LightSample ls = sampleLight( GetLight( lv, 0 ), surf.position, lp, p.polyLightSpotlightFactor );
        const float       tp = targetPdf( surf, ls, loose );
        updateReservoir( dir, 0, tp, 1.0f, rnd16( seed, salt++ ) );
    }
    normalizeReservoir( dir, 1 );
}

__device__ __forceinline__ Reservoir CombineInitial( Reservoir regular, const Reservoir& dir, uint32_t seed, uint32_t& salt )
{
    const float rnd = rnd16( seed, salt++ );
    updateCombined( regular, dir, rnd );
    normalizeReservoir( regular, 1 );
    return regular;
}

// traceShadowRay (RaygenCommon.h:322-350): from the surface biased toward
// the viewer to the light point, tmin 0.001, tmax = distance - 0.01
__device__ __forceinline__ RgExtRay ShadowRay( const Surface& s, v3 lightPos )
{
    const v3    start = s.position + s.toViewerDir * 0.01f;
    v3          l     = lightPos - start;
    const float dist  = length( l );
    return MakeRay( start, l / dist, 0.001f, dist - 0.01f );
}

// Cull masks of the ray types from the origin surface's instance custom
// index (RaygenCommon.h:103-177).
__host__ __device__ __forceinline__ uint32_t PrimaryCullMask( const RgExtFrameParams& p )
{
    return p.rayCullMaskWorld | IMASK_REFRACT | IMASK_FIRST_PERSON;
}
__device__ __forceinline__ uint32_t ReflRefrCullMask( const RgExtFrameParams& p, uint32_t custom, uint32_t geomFlags, bool refraction )
{
    uint32_t world = p.rayCullMaskWorld | IMASK_REFRACT;
    if( geomFlags & RG_MESH_PRIMITIVE_IGNORE_REFRACT_AFTER )
    {
        world &= ~IMASK_REFRACT;
        if( custom & CUSTOM_FIRST_PERSON )
        {
            return world;
        }
    }
    if( custom & CUSTOM_FIRST_PERSON )
    {
        return world | IMASK_FIRST_PERSON;
    }
    return world | ( refraction ? IMASK_FIRST_PERSON : IMASK_FIRST_PERSON_VIEWER );
}
// getShadowCullMask, and traceVisibility's ignoreFirstPersonViewer for the
// light lightIndexIgnoreFPVShadows (RaygenCommon.h:350-358)
__device__ __forceinline__ uint32_t ShadowCullMask( const RgExtFrameParams& p, uint32_t custom, uint32_t lightIndex )
{
    uint32_t m = p.rayCullMaskWorldShadow | ( ( custom & CUSTOM_FIRST_PERSON ) ? IMASK_FIRST_PERSON : IMASK_FIRST_PERSON_VIEWER );
    if( lightIndex == p.lightIndexIgnoreFPVShadows )
    {
        m &= ~IMASK_FIRST_PERSON_VIEWER;
    }
    return m;
}
__device__ __forceinline__ uint32_t IndirectCullMask( const RgExtFrameParams& p, uint32_t custom )
{
    return p.rayCullMaskWorld | ( ( custom & CUSTOM_FIRST_PERSON ) ? IMASK_FIRST_PERSON : IMASK_FIRST_PERSON_VIEWER );
}

// instanceCustomIndex of a hit's instance
__device__ __forceinline__ uint32_t HitCustomIndex( const RgExtFrameParams& p, const SceneView& sc, uint32_t prim )
{
    return CustomIndexOf( sc.geoms[ sc.triGeom[ prim ] ].flags, p.skyGeometry != 0u );
}

// doesPayloadContainHitInfo (RaygenCommon.h:191-205): a miss, or a hit on an
// instance with the SKY custom-index flag, reads as the sky.
__device__ __forceinline__ bool HitIsSky( const RgExtFrameParams& p, const SceneView& sc, const RgExtHit& h )
{
    return h.primitiveId == RG_EXT_MISS || ( p.skyGeometry != 0u && ( HitCustomIndex( p, sc, h.primitiveId ) & CUSTOM_SKY ) != 0u );
}

__device__ __forceinline__ LightPoint LightPointRnd( uint32_t seed )
{
    float u, v;
    rnd16_2( seed, SALT_LIGHT_POINT, &u, &v );
    return MakeLightPoint( u * 0.99f, v * 0.99f );
}

__device__ __forceinline__ bool DirectValid( const RgExtFrameParams& p, int bounce )
{
    return ( uint32_t( bounce ) < p.maxBounceShadowsLights || p.maxBounceShadowsLights == 0 ) &&
           ( p.lightCount + p.directionalLightExists > 0 );
}

// ------------------------------------------------------------ kernels

__global__ __launch_bounds__( THREADS ) void k_blue_noise( uint32_t* table )
{
    const uint32_t i = blockIdx.x * THREADS + threadIdx.x;
    if( i < 128u * 128u * 128u )
    {
        // SplitMix64 of the texel index: a white-noise stand-in for the
        // blue-noise KTX2 array the reference loads (BlueNoise.cpp:37-180)
        uint64_t z = 0x9E3779B97F4A7C15ull * ( uint64_t( i ) + 1ull ) + 0xC0FFEEull;
        z          = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
        z          = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
        table[ i ] = uint32_t( z ^ ( z >> 31 ) );
    }
}

// Primary rays over regular pixels (RaygenPrimary.inl:300-314).
__global__ __launch_bounds__( THREADS ) void k_primary_rays( RgExtFrameParams p, RgExtRay* rays )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const float u = ( float( x ) + 0.5f + p.jitterX ) / float( W );
    const float v = ( float( y ) + 0.5f + p.jitterY ) / float( H );
    const v3    d = GetRayDir( p, u, v );
    rays[ size_t( y ) * W + x ] = MakeRay( mk3( p.cameraPosition[ 0 ], p.cameraPosition[ 1 ], p.cameraPosition[ 2 ] ), d, p.primaryRayMinDist, p.rayLength );
}

// Moller-Trumbore barycentrics of a ray against a triangle's plane
// (intersectRayTriangle, HitInfo.inl:181-198)
__device__ __forceinline__ v3 RayTriBary( v3 p0, v3 p1, v3 p2, v3 o, v3 d )
{
    const v3    e1  = p1 - p0, e2 = p2 - p0;
    const v3    pv  = cross( d, e2 );
    const float inv = 1.0f / dot( e1, pv );
    const v3    tv  = o - p0;
    const v3    qv  = cross( tv, e1 );
    const float u   = dot( tv, pv ) * inv;
    const float v   = dot( d, qv ) * inv;
    return mk3( 1.0f - u - v, u, v );
}

// Media of the reflection / refraction chain and of the camera.
constexpr uint32_t RR_ACTIVE = 1u, RR_SPLIT = 2u, RR_DOSPLIT = 32u;
constexpr uint32_t MEDIA_VACUUM = 0u, MEDIA_WATER = 1u, MEDIA_GLASS = 2u, MEDIA_ACID = 3u;

// getMediaTypeFromFlags / isRefractFromFlags / isReflectFromFlags
// (Media.h:83-124) over RgMeshPrimitiveFlags as GeomInfoManager.cpp:131-186
// maps them to the instance flags.
__device__ __forceinline__ uint32_t MediaFromFlags( uint32_t f, float roughness )
{
    if( f & RG_MESH_PRIMITIVE_WATER )
    {
        return MEDIA_WATER;
    }
    if( ( f & RG_MESH_PRIMITIVE_GLASS ) || ( ( f & RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH ) && roughness < 0.1f ) )
    {
        return MEDIA_GLASS;
    }
    if( f & RG_MESH_PRIMITIVE_ACID )
    {
        return MEDIA_ACID;
    }
    return MEDIA_VACUUM;
}
__device__ __forceinline__ bool IsRefractFlags( uint32_t f, float roughness )
{
    return ( f & ( RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_ACID | RG_MESH_PRIMITIVE_GLASS ) ) ||
           ( ( f & RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH ) && roughness < 0.1f );
}
__device__ __forceinline__ bool IsReflectFlags( uint32_t f, float roughness )
{
    return ( f & ( RG_MESH_PRIMITIVE_MIRROR | RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_ACID | RG_MESH_PRIMITIVE_GLASS ) ) ||
           ( ( f & RG_MESH_PRIMITIVE_MIRROR_IF_SMOOTH ) && roughness < 0.1f );
}

// getIndexOfRefraction (Media.h:28-41)
__device__ __forceinline__ float IndexOfRefraction( const RgExtFrameParams& p, uint32_t media )
{
    return media == MEDIA_WATER || media == MEDIA_ACID ? p.indexOfRefractionWater : ( media == MEDIA_GLASS ? p.indexOfRefractionGlass : 1.0f );
}

// getMediaTransmittance (Media.h:43-58): exp(-distance * -log(colour)),
// acid extinction scaled by max(1, sqrt(density)).  Vacuum and glass do not
// attenuate (also for an infinite distance).
__device__ v3 MediaTransmittance( const RgExtFrameParams& p, uint32_t media, float distance )
{
    if( media != MEDIA_WATER && media != MEDIA_ACID )
    {
        return mk3( 1.0f, 1.0f, 1.0f );
    }
    const float* c = media == MEDIA_WATER ? p.waterColorAndDensity : p.acidColorAndDensity;
    v3           e = mk3( -logf( c[ 0 ] ), -logf( c[ 1 ] ), -logf( c[ 2 ] ) );
    if( media == MEDIA_ACID )
    {
        e = e * fmaxf( 1.0f, sqrtf( c[ 3 ] ) );
    }
    return mk3( expf( -distance * e.x ), expf( -distance * e.y ), expf( -distance * e.z ) );
}

// calcRefractionDirection, Ray Tracing Gems II ch. 8 (Media.h:62-77):
// false on total internal reflection.
__device__ __forceinline__ bool RefractDir( float n1, float n2, v3 I, v3 N, v3* T )
{
    const float eta = n1 / n2;
    const float c1  = -dot( I, N );
    const float w   = eta * c1;
    const float c2m = ( w - eta ) * ( w + eta );
    if( c2m < -1.0f )
    {
        return false;
    }
    *T = I * eta + N * ( w - sqrtf( 1.0f + c2m ) );
    return true;
}

__device__ __forceinline__ float Pow5( float x )
{
    const float x2 = x * x;
    return x2 * x2 * x;
}

// getFresnelSchlick(n1, n2, V, N) (BRDF.h:78-84)
__device__ __forceinline__ float FresnelIor( float n1, float n2, v3 V, v3 N )
{
    float r0 = ( n1 - n2 ) / ( n1 + n2 );
    r0 *= r0;
    const float t = Pow5( 1.0f - fabsf( dot( N, V ) ) );
    return r0 * ( 1.0f - t ) + t;
}

// getFresnelSchlick(nl, specularColor) (BRDF.h:73-76)
__device__ __forceinline__ v3 FresnelSpec( float nl, v3 f0 )
{
    const float t = Pow5( 1.0f - fmaxf( nl, 0.0f ) );
    return f0 + ( mk3( 1.0f, 1.0f, 1.0f ) - f0 ) * t;
}

// getGlowingMediaFog (Media.h:63-77): acid glows with distance
__device__ v3 GlowingMediaFog( const RgExtFrameParams& p, uint32_t media, float distance )
{
    if( media != MEDIA_ACID )
    {
        return mk3( 0.0f, 0.0f, 0.0f );
    }
    const float density = 0.00005f * p.acidColorAndDensity[ 3 ];
    const float fog     = clampf( 1.0f - expf( -distance * density ), 0.0f, 1.0f );
    return mk3( p.acidColorAndDensity[ 0 ], p.acidColorAndDensity[ 1 ], p.acidColorAndDensity[ 2 ] ) * fog;
}

// storeSky (RaygenPrimary.inl:90-139) for a colour sky; the motion of an
// infinitely far point uses the checkerboard pixel, as :74-76 and :118 do.
// The primary pass also clears DEPTH_GRAD; reflections keep it.
__device__ void StoreSky( const RgExtFrameParams& p, const FbPtrs& P, size_t ci, size_t ri, int cx, int cy, v3 dir, v3 thr, float thrW, bool primary )
{
    const int   W  = p.renderWidth, H = p.renderHeight;
    const float cu = ( float( cx ) + 0.5f + p.jitterX ) / float( W );
    const float cv = ( float( cy ) + 0.5f + p.jitterY ) / float( H );
    const v3    rd = GetRayDir( p, cu, cv );
    float       w;
    const v3    vc = MulM4( p.view, rd, 0.0f, &w ), vp = MulM4( p.viewPrev, rd, 0.0f, &w );
    // mat3(projection) * v: the upper-left 3x3 block, no division by w
    const v3 cc = mk3( p.projection[ 0 ] * vc.x + p.projection[ 4 ] * vc.y + p.projection[ 8 ] * vc.z,
                       p.projection[ 1 ] * vc.x + p.projection[ 5 ] * vc.y + p.projection[ 9 ] * vc.z, 0.0f );
    const v3 cp = mk3( p.projectionPrev[ 0 ] * vp.x + p.projectionPrev[ 4 ] * vp.y + p.projectionPrev[ 8 ] * vp.z,
                       p.projectionPrev[ 1 ] * vp.x + p.projectionPrev[ 5 ] * vp.y + p.projectionPrev[ 9 ] * vp.z, 0.0f );
    const float mx  = ( cp.x * 0.5f + 0.5f ) - ( cc.x * 0.5f + 0.5f );
    const float my  = ( cp.y * 0.5f + 0.5f ) - ( cc.y * 0.5f + 0.5f );
    const v3    sky = SkyColor( p );
    Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ ci ]           = 1u;
    Plane< float4 >( P, RG_EXT_PLANE_ALBEDO )[ ri ]             = make_float4( sky.x, sky.y, sky.z, 0.0f );
    Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ ci ]           = EncodeNormal( mk3( 0, 0, 0 ) );
    Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ ci ] = make_float2( 0.f, 0.f );
    Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ ci ]         = MAX_RAY_LEN * 2.0f;
    Plane< float4 >( P, RG_EXT_PLANE_MOTION )[ ci ]             = make_float4( mx, my, 0.f, 0.f );
    Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION )[ ci ]   = make_float4( 1e7f, 1e7f, 1e7f, 1e7f );
    Plane< float4 >( P, RG_EXT_PLANE_VISIBILITY_BUFFER )[ ci ]  = make_float4( 4294967295.0f, 4294967295.0f, 4294967295.0f, 4294967295.0f );
    Plane< float4 >( P, RG_EXT_PLANE_VIEW_DIRECTION )[ ci ]     = make_float4( dir.x, dir.y, dir.z, 0.f );
    Plane< float4 >( P, RG_EXT_PLANE_THROUGHPUT )[ ci ]         = make_float4( thr.x, thr.y, thr.z, thrW );
    Plane< float4 >( P, RG_EXT_PLANE_SCREEN_EMISSION )[ ri ]    = make_float4( 0.f, 0.f, 0.f, 0.f );
    Plane< float4 >( P, RG_EXT_PLANE_ACID_FOG )[ ri ]           = make_float4( 0.f, 0.f, 0.f, 0.f );
    if( primary )
    {
        Plane< float >( P, RG_EXT_PLANE_DEPTH_GRAD )[ ci ] = 0.0f;
    }
}

// G-buffer of the primary hit (RaygenPrimary.inl:280-348, storeSky :90-139).
template< bool TEX >
__global__ __launch_bounds__( THREADS ) void k_primary_shade( RgExtFrameParams p, const RgExtHit* hits, SceneView sc, FbPtrs P )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    int cx, cy;
    cbFromRegular( x, y, W, &cx, &cy );
    const size_t   ci  = size_t( cy ) * W + cx;
    const size_t   ri  = size_t( y ) * W + x;
    const RgExtHit hit = hits[ ri ];
    const float    u   = ( float( x ) + 0.5f + p.jitterX ) / float( W );
    const float    v   = ( float( y ) + 0.5f + p.jitterY ) / float( H );
    const v3       dir = GetRayDir( p, u, v );
    const v3       org = mk3( p.cameraPosition[ 0 ], p.cameraPosition[ 1 ], p.cameraPosition[ 2 ] );

    if( HitIsSky( p, sc, hit ) )
    {
        StoreSky( p, P, ci, ri, cx, cy, dir, mk3( 1.f, 1.f, 1.f ), 0.0f, true );
        return;
    }

    const float*      t  = sc.tris + size_t( hit.primitiveId ) * 9;
    const v3 p0 = mk3( t[ 0 ], t[ 1 ], t[ 2 ] ), p1 = mk3( t[ 3 ], t[ 4 ], t[ 5 ] ), p2 = mk3( t[ 6 ], t[ 7 ], t[ 8 ] );

    // ray differentials on the same triangle (HitInfo.inl:243-268); their
    // barycentric deltas give the texture gradients
    const v3 dAX = GetRayDir( p, u + 1.0f / float( W ), v );
    const v3 dAY = GetRayDir( p, u, v + 1.0f / float( H ) );
    const v3 bAX = RayTriBary( p0, p1, p2, org, dAX );
    const v3 bAY = RayTriBary( p0, p1, p2, org, dAY );
    TexQuery tq;
    tq.mode = TQ_GRAD;
    tq.dbx  = bAX - mk3( 1.0f - hit.u - hit.v, hit.u, hit.v );
    tq.dby  = bAY - mk3( 1.0f - hit.u - hit.v, hit.u, hit.v );
    const HitInfo     h  = GetHit< TEX >( sc, p, hit.primitiveId, hit.u, hit.v, org, tq );
    const GeomRecord& gr = sc.geoms[ h.geom ];
    const v3 prevPos = gr.isStatic ? h.position : Xform34( gr.prevFromCur, h.position );

    float    wc, wp, wx, wy;
    const v3 vCur  = MulM4( p.view, h.position, 1.0f, &wc );
    const v3 vPrev = MulM4( p.viewPrev, prevPos, 1.0f, &wp );
    const v3 vAX   = MulM4( p.view, p0 * bAX.x + p1 * bAX.y + p2 * bAX.z, 1.0f, &wx );
    const v3 vAY   = MulM4( p.view, p0 * bAY.x + p1 * bAY.y + p2 * bAY.z, 1.0f, &wy );
    float    cwc, cwp;
    const v3 cCur  = MulM4( p.projection, vCur, wc, &cwc );
    const v3 cPrev = MulM4( p.projectionPrev, vPrev, wp, &cwp );
    // dot(projection[2], v): the third column of the matrix, as the reference writes it
    const float* col2 = p.projection + 8;
    const float  dAXc = col2[ 0 ] * vAX.x + col2[ 1 ] * vAX.y + col2[ 2 ] * vAX.z + col2[ 3 ] * wx;
    const float  dAYc = col2[ 0 ] * vAY.x + col2[ 1 ] * vAY.y + col2[ 2 ] * vAY.z + col2[ 3 ] * wy;
    const float  gx = dAXc - cCur.z, gy = dAYc - cCur.z;

    const float sxc = ( cCur.x / cwc ) * 0.5f + 0.5f, syc = ( cCur.y / cwc ) * 0.5f + 0.5f;
    const float sxp = ( cPrev.x / cwp ) * 0.5f + 0.5f, syp = ( cPrev.y / cwp ) * 0.5f + 0.5f;
    const float depthLinear = length( vCur );
    const float motionZ     = length( vPrev ) - depthLinear;

    Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ ci ]           = 0u;
    Plane< float4 >( P, RG_EXT_PLANE_ALBEDO )[ ri ]             = make_float4( h.albedo.x, h.albedo.y, h.albedo.z, 0.0f );
    Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ ci ]           = EncodeNormal( h.normal );
    Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ ci ] = make_float2( h.metallic, h.roughness );
    Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ ci ]         = depthLinear;
    Plane< float >( P, RG_EXT_PLANE_DEPTH_GRAD )[ ci ]          = sqrtf( gx * gx + gy * gy );
    Plane< float4 >( P, RG_EXT_PLANE_MOTION )[ ci ]             = make_float4( sxp - sxc, syp - syc, motionZ, 0.0f );
    Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION )[ ci ] =
        make_float4( h.position.x, h.position.y, h.position.z, __uint_as_float( CustomIndexOf( gr.flags, p.skyGeometry != 0u ) ) );
    Plane< uint4 >( P, RG_EXT_PLANE_VISIBILITY_BUFFER )[ ci ]   = make_uint4( 1u, ( h.localPrim << 12 ) | ( h.geom & 0xFFFu ), __float_as_uint( hit.u ), __float_as_uint( hit.v ) );
    Plane< float4 >( P, RG_EXT_PLANE_VIEW_DIRECTION )[ ci ]     = make_float4( dir.x, dir.y, dir.z, 0.0f );
    // the camera's medium attenuates the primary segment (RaygenPrimary.inl:321-328)
    const v3 thr  = MediaTransmittance( p, p.cameraMediaType & 3u, depthLinear );
    const v3 emis = h.emission * thr;
    const v3 fog  = GlowingMediaFog( p, p.cameraMediaType & 3u, depthLinear );
    Plane< float4 >( P, RG_EXT_PLANE_THROUGHPUT )[ ci ]         = make_float4( thr.x, thr.y, thr.z, 0.f );
    Plane< float4 >( P, RG_EXT_PLANE_SCREEN_EMISSION )[ ri ]    = make_float4( emis.x, emis.y, emis.z, 0.f );
    Plane< float4 >( P, RG_EXT_PLANE_ACID_FOG )[ ri ]           = make_float4( fog.x, fog.y, fog.z, 0.f );
}

// ------------------------------------------- reflection / refraction chain
// RAYGEN_REFL_REFR_SHADER (RaygenPrimary.inl:352-599) as a wavefront: each
// depth is k_reflrefr_a (decide, build the ray), one closest-hit launch, then
// k_reflrefr_b (store the hit or the sky into the G-buffer).  The loop's
// local state lives in three float4 per pixel:
//   S[0] = throughput.xyz, bits (1: still tracing, 2: wasSplit, media << 2)
//   S[1] = ray direction.xyz, full path length
//   S[2] = virtual position for motion.xyz
//   S[3] = .x: the ray-cone width the last hit was shaded with
//   S[4] = screen emission.xyz, S[5] = acid fog.xyz (framebufScreenEmisRT,
//          framebufAcidFogRT of the loop)
// Portals, animated water normals, screen emission and acid fog are not
// part of this path (they need the water texture or portal instances).
struct RrState
{
    float4* s; // 3 x W*H float4
    size_t  n;
};

// One depth of the loop up to traceReflectionRefractionRay.  Depth 0 reads
// the primary surface from the G-buffer (as the reference restores it) and
// its flags from the primary hit; later depths recompute the last hit.
template< bool TEX >
__global__ __launch_bounds__( THREADS ) void k_reflrefr_a( RgExtFrameParams p, FbPtrs P, SceneView sc, RgExtRay* rays, uint8_t* masks,
                                                           const RgExtHit* hits, RrState st, uint32_t depth )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    int cx, cy;
    cbFromRegular( x, y, W, &cx, &cy );
    const size_t ci = size_t( cy ) * W + cx;
    const size_t ri = size_t( y ) * W + x;

    v3       pos, normal, albedo, thr, dir;
    float    roughness, metallic, fpl;
    uint32_t flags, bits, custom;
    if( depth == 0 )
    {
        if( Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ ci ] != 0u )
        {
            st.s[ ri ]  = make_float4( 0.f, 0.f, 0.f, __uint_as_float( 0u ) );
            rays[ ri ]  = NullRay();
            return;
        }
        const float4 alb = Plane< float4 >( P, RG_EXT_PLANE_ALBEDO )[ ri ];
        const float4 sp  = Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION )[ ci ];
        const float2 mr  = Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ ci ];
        const float4 th  = Plane< float4 >( P, RG_EXT_PLANE_THROUGHPUT )[ ci ];
        albedo           = mk3( alb.x, alb.y, alb.z );
        pos              = mk3( sp.x, sp.y, sp.z );
        normal           = DecodeNormal( Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ ci ] );
        metallic         = mr.x;
        roughness        = mr.y;
        thr              = mk3( th.x, th.y, th.z );
        const float u    = ( float( x ) + 0.5f + p.jitterX ) / float( W );
        const float v    = ( float( y ) + 0.5f + p.jitterY ) / float( H );
        dir              = GetRayDir( p, u, v );
        fpl              = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ ci ];
        flags            = sc.geoms[ sc.triGeom[ hits[ ri ].primitiveId ] ].flags;
        custom           = __float_as_uint( sp.w );
        bits             = RR_ACTIVE | ( ( p.cameraMediaType & 3u ) << 2 );
        // ray cone (RaygenPrimary.inl:396-406): width 0, then propagated
        // over the primary hit's distance
        st.s[ 2 * st.n + ri ] = make_float4( pos.x, pos.y, pos.z, p.cameraRayConeSpreadAngle * fpl );
        st.s[ 4 * st.n + ri ] = Plane< float4 >( P, RG_EXT_PLANE_SCREEN_EMISSION )[ ri ];
        st.s[ 5 * st.n + ri ] = Plane< float4 >( P, RG_EXT_PLANE_ACID_FOG )[ ri ];
    }
    else
    {
        const float4 s0 = st.s[ ri ];
        bits            = __float_as_uint( s0.w );
        if( ( bits & RR_ACTIVE ) == 0u )
        {
            rays[ ri ] = NullRay();
            return;
        }
        const float4   s1  = st.s[ st.n + ri ];
        const RgExtHit hit = hits[ ri ];
        const RgExtRay r   = rays[ ri ];
        TexQuery       cq;
        cq.mode            = TQ_CONE;
        cq.a               = st.s[ 3 * st.n + ri ].x; // the cone the last hit was shaded with
        cq.dir             = mk3( s1.x, s1.y, s1.z );
        const HitInfo  h   = GetHit< TEX >( sc, p, hit.primitiveId, hit.u, hit.v, mk3( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] ), cq );
        pos                = h.position;
        normal             = h.normal;
        albedo             = h.albedo;
        metallic           = h.metallic;
        roughness          = h.roughness;
        flags              = sc.geoms[ h.geom ].flags;
        custom             = CustomIndexOf( flags, p.skyGeometry != 0u );
        thr                = mk3( s0.x, s0.y, s0.z );
        dir                = mk3( s1.x, s1.y, s1.z );
        fpl                = s1.w;
    }

    // getNewRayMedia (RaygenPrimary.inl:141-150)
    const uint32_t curMedia = ( bits >> 2 ) & 3u;
    uint32_t       newMedia = ( depth == 0 && p.cameraMediaType != MEDIA_VACUUM ) ? MEDIA_VACUUM : MediaFromFlags( flags, roughness );
    const bool     toRefract = IsRefractFlags( flags, roughness );
    const bool     toReflect = IsReflectFlags( flags, roughness );
    if( !toReflect && !toRefract )
    {
        st.s[ ri ] = make_float4( thr.x, thr.y, thr.z, __uint_as_float( bits & ~RR_ACTIVE ) );
        rays[ ri ] = NullRay();
        return;
    }
    const float curIor  = IndexOfRefraction( p, curMedia );
    const float newIor  = IndexOfRefraction( p, newMedia );
    const bool  isWater = newMedia == MEDIA_WATER || curMedia == MEDIA_WATER || newMedia == MEDIA_ACID || curMedia == MEDIA_ACID;
    // getNormal (RaygenPrimary.inl:246-270): face the ray; water and acid
    // take getWaterNormal (:152-199) with the reference's default water
    // texture, one texel (127, 127, 255): every octave reads
    // (-1/255, -1/255, 1), so the normal is the basis of getONB (Random.h:
    // 160-169) times normalize((0,0,1) + strength * 0.55 * (-1/255, -1/255, 1))
    if( dot( normal, -dir ) < 0.0f )
    {
        normal = -normal;
    }
    if( isWater )
    {
        v3 b0, b1;
        revisedONB( normal, &b0, &b1 );
        const float k  = p.waterWaveStrength * 0.55f;
        const v3    tn = normalize( mk3( k * ( -1.0f / 255.0f ), k * ( -1.0f / 255.0f ), 1.0f + k ) );
        normal         = b0 * tn.x + b1 * tn.y + normal * tn.z;
    }
    else
    {
        normal = normalize( normal );
    }

    v3       origin  = pos;
    bool     wasSplit = ( bits & RR_SPLIT ) != 0u;
    bool     doSplit  = !wasSplit;
    bool     doRefr   = false;
    float    F        = 1.0f;
    v3       refr;
    uint32_t media = curMedia;
    if( toRefract && RefractDir( curIor, newIor, dir, normal, &refr ) )
    {
        doRefr = cx >= cbSeparator( W ); // isCheckerboardPixOdd
        F      = FresnelIor( curIor, newIor, -dir, normal );
    }
    else
    {
        doSplit = false; // total internal reflection
    }
    if( doRefr )
    {
        dir = refr;
        thr = thr * ( 1.0f - F );
        if( flags & RG_MESH_PRIMITIVE_THIN_MEDIA )
        {
            // a slab of thinMediaWidth along the normal, then back out
            const float len = p.thinMediaWidth / fmaxf( 0.001f, -dot( normal, dir ) );
            origin          = origin + dir * len;
            thr             = thr * MediaTransmittance( p, newMedia, len );
            fpl += len;
            if( RefractDir( newIor, curIor, dir, normal, &refr ) )
            {
                dir = refr;
            }
            newMedia = curMedia;
        }
        media = newMedia;
    }
    else
    {
        // reflect(normalize(I), N)
        const v3 i = normalize( dir );
        dir        = i - normal * ( 2.0f * dot( normal, i ) );
        if( !isWater )
        {
            thr = thr * FresnelSpec( fmaxf( 0.0f, dot( normal, dir ) ), SpecularColor( albedo, metallic ) );
        }
        else
        {
            thr    = thr * F;
            origin = origin + dir * 0.005f; // SHIPPING_HACK offset of the reference
        }
    }
    if( doSplit )
    {
        thr      = thr * 2.0f;
        wasSplit = true;
    }
    bits = RR_ACTIVE | ( wasSplit ? RR_SPLIT : 0u ) | ( doSplit ? RR_DOSPLIT : 0u ) | ( ( media & 3u ) << 2 );
    st.s[ ri ]        = make_float4( thr.x, thr.y, thr.z, __uint_as_float( bits ) );
    st.s[ st.n + ri ] = make_float4( dir.x, dir.y, dir.z, fpl );
    // traceReflectionRefractionRay (RaygenCommon.h:235-251)
    rays[ ri ]  = MakeRay( origin, dir, 0.001f, p.rayLength );
    masks[ ri ] = uint8_t( ReflRefrCullMask( p, custom, flags, doRefr ) );
}

// After the trace: the sky (RaygenPrimary.inl:550-556) or the new surface
// (getHitInfoWithRayCone_ReflectionRefraction, HitInfo.inl:213-330, and the
// G-buffer stores of :584-597), with its virtual-position motion.
template< bool TEX >
__global__ __launch_bounds__( THREADS ) void k_reflrefr_b( RgExtFrameParams p, FbPtrs P, SceneView sc, const RgExtRay* rays, const RgExtHit* hits,
                                                           RrState st, uint32_t depth )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t ri = size_t( y ) * W + x;
    const float4 s0 = st.s[ ri ];
    uint32_t     bits = __float_as_uint( s0.w );
    if( ( bits & RR_ACTIVE ) == 0u )
    {
        return;
    }
    int cx, cy;
    cbFromRegular( x, y, W, &cx, &cy );
    const size_t   ci    = size_t( cy ) * W + cx;
    const float4   s1    = st.s[ st.n + ri ];
    const v3       dir   = mk3( s1.x, s1.y, s1.z );
    const uint32_t media = ( bits >> 2 ) & 3u;
    const float    thrW  = ( bits & RR_SPLIT ) ? 1.0f : -1.0f;
    v3             thr   = mk3( s0.x, s0.y, s0.z );
    const RgExtHit hit   = hits[ ri ];
    if( HitIsSky( p, sc, hit ) )
    {
        const float up = fabsf( dot( dir, mk3( p.worldUp[ 0 ], p.worldUp[ 1 ], p.worldUp[ 2 ] ) ) );
        thr            = thr * MediaTransmittance( p, media, 1.0f / ( up * up * up ) );
        StoreSky( p, P, ci, ri, cx, cy, dir, thr, thrW, false );
        st.s[ ri ] = make_float4( thr.x, thr.y, thr.z, __uint_as_float( bits & ~RR_ACTIVE ) );
        return;
    }
    const RgExtRay r      = rays[ ri ];
    const v3       origin = mk3( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] );
    const float4   s2     = st.s[ 2 * st.n + ri ];
    TexQuery       cq;
    cq.mode               = TQ_CONE;
    cq.a                  = s2.w;
    cq.dir                = dir;
    const HitInfo  h      = GetHit< TEX >( sc, p, hit.primitiveId, hit.u, hit.v, origin, cq );
    const float    rayLen = length( h.position - origin );
    // propagateRayCone (RayCone.h:30-36): the spread angle doubles at every
    // propagation, so at this depth it is the camera's times 2^(depth + 1)
    const float spread    = p.cameraRayConeSpreadAngle * exp2f( float( depth + 1 ) );

    // virtualPosForMotion += viewDir * rayLen, seen by both cameras
    const float u   = ( float( x ) + 0.5f + p.jitterX ) / float( W );
    const float v   = ( float( y ) + 0.5f + p.jitterY ) / float( H );
    const v3    cam = GetRayDir( p, u, v );
    const v3    vp  = mk3( s2.x, s2.y, s2.z ) + cam * rayLen;
    float       wc, wp, cwc, cwp;
    const v3    vCur  = MulM4( p.view, vp, 1.0f, &wc );
    const v3    vPrev = MulM4( p.viewPrev, vp, 1.0f, &wp );
    const v3    cCur  = MulM4( p.projection, vCur, wc, &cwc );
    const v3    cPrev = MulM4( p.projectionPrev, vPrev, wp, &cwp );
    const float sxc = ( cCur.x / cwc ) * 0.5f + 0.5f, syc = ( cCur.y / cwc ) * 0.5f + 0.5f;
    const float sxp = ( cPrev.x / cwp ) * 0.5f + 0.5f, syp = ( cPrev.y / cwp ) * 0.5f + 0.5f;
    const float motionZ = length( vPrev ) - length( vCur );

    thr             = thr * MediaTransmittance( p, media, rayLen );
    const float fpl = s1.w + rayLen;
    // screen emission and acid fog along the chain (RaygenPrimary.inl:575-589)
    const float4 e4   = st.s[ 4 * st.n + ri ], f4 = st.s[ 5 * st.n + ri ];
    const v3     emis = mk3( e4.x, e4.y, e4.z ) + h.emission * thr;
    const v3     fog  = mk3( f4.x, f4.y, f4.z ) + GlowingMediaFog( p, media, rayLen ) * ( ( bits & RR_DOSPLIT ) ? 2.0f : 1.0f );
    const v3     semi = p.cameraMediaType != MEDIA_ACID ? emis + fog * 0.05f : emis;
    st.s[ 4 * st.n + ri ] = make_float4( emis.x, emis.y, emis.z, 0.0f );
    st.s[ 5 * st.n + ri ] = make_float4( fog.x, fog.y, fog.z, 0.0f );
    Plane< float4 >( P, RG_EXT_PLANE_SCREEN_EMISSION )[ ri ] = make_float4( semi.x, semi.y, semi.z, 0.0f );
    Plane< float4 >( P, RG_EXT_PLANE_ACID_FOG )[ ri ]        = make_float4( fog.x, fog.y, fog.z, 0.0f );

    Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ ci ]           = 0u;
    Plane< float4 >( P, RG_EXT_PLANE_ALBEDO )[ ri ]             = make_float4( h.albedo.x, h.albedo.y, h.albedo.z, 0.0f );
    Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ ci ]           = EncodeNormal( h.normal );
    Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ ci ] = make_float2( h.metallic, h.roughness );
    Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ ci ]         = fpl;
    Plane< float4 >( P, RG_EXT_PLANE_MOTION )[ ci ]             = make_float4( sxp - sxc, syp - syc, motionZ, 0.0f );
    Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION )[ ci ] =
        make_float4( h.position.x, h.position.y, h.position.z, __uint_as_float( CustomIndexOf( sc.geoms[ h.geom ].flags, p.skyGeometry != 0u ) ) );
    Plane< uint4 >( P, RG_EXT_PLANE_VISIBILITY_BUFFER )[ ci ]   = make_uint4( 1u, ( h.localPrim << 12 ) | ( h.geom & 0xFFFu ), __float_as_uint( hit.u ), __float_as_uint( hit.v ) );
    Plane< float4 >( P, RG_EXT_PLANE_VIEW_DIRECTION )[ ci ]     = make_float4( dir.x, dir.y, dir.z, 0.0f );
    Plane< float4 >( P, RG_EXT_PLANE_THROUGHPUT )[ ci ]         = make_float4( thr.x, thr.y, thr.z, thrW );
    st.s[ ri ]                 = make_float4( thr.x, thr.y, thr.z, __uint_as_float( bits ) );
    st.s[ st.n + ri ]          = make_float4( dir.x, dir.y, dir.z, fpl );
    st.s[ 2 * st.n + ri ]      = make_float4( vp.x, vp.y, vp.z, s2.w + spread * rayLen );
    st.s[ 3 * st.n + ri ]      = make_float4( s2.w, 0.0f, 0.0f, 0.0f );
}

// RtInitialReservoirs.rgen, first half: candidates + directional visibility ray.
__global__ __launch_bounds__( THREADS ) void k_initial_a( RgExtFrameParams p, FbPtrs P, LightView lv, RgExtRay* rays, uint8_t* masks, uint4* pend )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t  i = size_t( y ) * W + x;
    const Surface s = FetchSurface( P, W, x, y );
    if( s.isSky )
    {
        Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS_INITIAL )[ i ] = packReservoir( emptyReservoir() );
        rays[ i ]                                                 = NullRay();
        pend[ i ]                                                 = make_uint4( 0u, 0u, 0u, 0u );
        return;
    }
    const uint32_t seed = randomSeed( x, y, p.frameId );
    uint32_t       salt = SALT_INITIAL_RESERVOIRS;
    const LightPoint lp = LightPointRnd( seed );
    Reservoir regular, dir;
    InitialCandidates( p, lv, s, seed, salt, lp, false, regular, dir );
    Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS_INITIAL )[ i ] = packReservoir( regular );
    const uint4 pd = packReservoir( dir );
    pend[ i ]      = make_uint4( pd.x, pd.y, pd.z, salt );
    if( isValid( dir ) )
    {
        const LightSample ls = sampleLight( GetLight( lv, 0 ), s.position, lp, p.polyLightSpotlightFactor );
        rays[ i ]            = ShadowRay( s, ls.position );
        masks[ i ]           = uint8_t( ShadowCullMask( p, s.instCustomIndex, 0u ) );
    }
    else
    {
        rays[ i ] = NullRay();
    }
}

// Second half: drop an occluded directional sample and combine.
__global__ __launch_bounds__( THREADS ) void k_initial_b( RgExtFrameParams p, FbPtrs P, const RgExtHit* hits, const uint4* pend )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t i = size_t( y ) * W + x;
    if( Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ i ] != 0u )
    {
        return;
    }
    const uint4 pd  = pend[ i ];
    Reservoir   dir = unpackReservoir( make_uint4( pd.x, pd.y, pd.z, 0u ) );
    if( isValid( dir ) && hits[ i ].primitiveId != RG_EXT_MISS )
    {
        dir = emptyReservoir();
    }
    uint32_t  salt    = pd.w;
    Reservoir regular = unpackReservoir( Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS_INITIAL )[ i ] );
    const uint32_t seed = randomSeed( x, y, p.frameId );
    Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS_INITIAL )[ i ] = packReservoir( CombineInitial( regular, dir, seed, salt ) );
}

// testSurfaceForReuse (RaygenCommon.h:488-500)
__device__ __forceinline__ bool ReuseOk( int lo, int hi, int H, int qx, int qy, float dCur, float dOther, v3 nCur, v3 nOther )
{
    return inRenderArea( qx, qy, lo, hi, H ) && ( fabsf( dCur - dOther ) / fabsf( dCur ) < 0.1f ) && ( dot( nCur, nOther ) > 0.5f );
}

template< typename T >
__device__ __forceinline__ T LdPlane( const T* base, int W, int H, int x, int y, T zero )
{
    return ( x < 0 || y < 0 || x >= W || y >= H ) ? zero : base[ size_t( y ) * W + x ];
}

// RtRaygenDirect.rgen first half: spatiotemporal reuse, shading, shadow ray.
__global__ __launch_bounds__( THREADS ) void k_direct_a( RgExtFrameParams p, FbPtrs P, LightView lv, RgExtRay* rays, uint8_t* masks, float4* pendD,
                                                         float4* pendS )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t  i = size_t( y ) * W + x;
    const Surface s = FetchSurface( P, W, x, y );
    rays[ i ]       = NullRay();
    if( s.isSky )
    {
        Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS )[ i ] = packReservoir( emptyReservoir() );
        pendD[ i ] = make_float4( 0.f, 0.f, 0.f, -1.0f ); // w < 0: sky
        pendS[ i ] = make_float4( 0.f, 0.f, 0.f, 0.f );
        return;
    }
    const uint32_t seed = randomSeed( x, y, p.frameId );
    float          dist = MAX_RAY_LEN;
    v3             diff = mk3( 0, 0, 0 ), spec = mk3( 0, 0, 0 );
    Reservoir      res  = emptyReservoir();
    if( DirectValid( p, 0 ) )
    {
        const LightPoint lp = LightPointRnd( seed );
        // selectLight_Direct (RaygenCommon.h:503-599)
        int lo, hi;
        cbRenderArea( x, W, &lo, &hi );
        const float4 mot      = Plane< float4 >( P, RG_EXT_PLANE_MOTION )[ i ];
        const float  depthCur = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ i ];
        const float  px = ( ( float( x ) + 0.5f ) * ( 1.0f / ( float( W ) / 2.0f ) ) + mot.x ) * ( float( W ) / 2.0f );
        const float  py = ( ( float( y ) + 0.5f ) * ( 1.0f / float( H ) ) + mot.y ) * float( H );
        uint32_t     salt = SALT_LIGHT_CHOOSE_DIRECT;
        const Reservoir init = unpackReservoir( Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS_INITIAL )[ i ] );
        Reservoir       comb = init;
        for( uint32_t k = 0; k < p.restirTemporalSamples; k++ )
        {
            float r4[ 4 ];
            rnd8_4( seed, salt++, r4 );
            const int   qx    = int( floorf( px + ( r4[ 0 ] * 2.0f - 1.0f ) * 2.0f ) );
            const int   qy    = int( floorf( py + ( r4[ 1 ] * 2.0f - 1.0f ) * 2.0f ) );
            // one round trip at the clamped pixel (an outside one fails ReuseOk)
            const size_t   qi    = size_t( min( max( qy, 0 ), H - 1 ) ) * W + min( max( qx, 0 ), W - 1 );
            const float    dPrev = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD, true )[ qi ];
            const uint32_t nPrev = Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL, true )[ qi ];
            const uint4    tRaw  = Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS, true )[ qi ];
            if( !ReuseOk( lo, hi, H, qx, qy, depthCur, dPrev - mot.z, s.normal, DecodeNormal( nPrev ) ) )
            {
                continue;
            }
            Reservoir t = unpackReservoir( tRaw );
            normalizeReservoir( t, init.M * 20 );
            float tpCur = 0.0f;
            if( t.selected != LIGHT_INDEX_NONE )
            {
                const uint32_t cur = t.selected < lv.prevCount ? lv.prevToCur[ t.selected ] : 0xFFFFFFFFu;
                if( cur != 0xFFFFFFFFu && cur != LIGHT_INDEX_NONE )
                {
                    tpCur      = targetPdf( s, sampleLight( GetLight( lv, cur ), s.position, lp, p.polyLightSpotlightFactor ), false );
                    t.selected = cur;
                }
            }
            updateCombinedNewSurf( comb, t, tpCur, rnd16( seed, salt++ ) );
        }
        for( uint32_t k = 0; k < p.restirSpatialSamples; k++ )
        {
            float r4[ 4 ];
            rnd8_4( seed, salt++, r4 );
            const int   qx = x + int( ( r4[ 0 ] * 2.0f - 1.0f ) * 30.0f );
            const int   qy = y + int( ( r4[ 1 ] * 2.0f - 1.0f ) * 30.0f );
            const size_t   qi  = size_t( min( max( qy, 0 ), H - 1 ) ) * W + min( max( qx, 0 ), W - 1 );
            const float    dO  = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ qi ];
            const uint32_t nO  = Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ qi ];
            const uint4    spR = Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS_INITIAL )[ qi ];
            if( !ReuseOk( lo, hi, H, qx, qy, depthCur, dO, s.normal, DecodeNormal( nO ) ) )
            {
                continue;
            }
            const Reservoir sp = unpackReservoir( spR );
            updateCombined( comb, sp, rnd16( seed, salt++ ) );
        }
        if( !( comb.weightSum <= 0.0f || comb.selected == LIGHT_INDEX_NONE ) )
        {
            res = comb;
            // traceDirectIllumination (RaygenCommon.h:603-637)
            const LightSample ls = sampleLight( GetLight( lv, res.selected ), s.position, lp, p.polyLightSpotlightFactor );
            shadeLight( s, ls, selectedSampleWeight( res ), false, &diff, &spec );
            if( luminance( diff + spec ) <= 0.0f )
            {
                diff = spec = mk3( 0, 0, 0 );
            }
            else if( 0u < p.maxBounceShadowsLights )
            {
                rays[ i ]  = ShadowRay( s, ls.position );
                masks[ i ] = uint8_t( ShadowCullMask( p, s.instCustomIndex, res.selected ) );
                dist       = length( ls.position - s.position );
            }
        }
    }
    Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS )[ i ] = packReservoir( res );
    pendD[ i ] = make_float4( diff.x, diff.y, diff.z, dist );
    pendS[ i ] = make_float4( spec.x, spec.y, spec.z, 0.0f );
}

// Second half: visibility, demodulation and the direct-light planes.
__global__ __launch_bounds__( THREADS ) void k_direct_b( RgExtFrameParams p, FbPtrs P, const RgExtRay* rays, const RgExtHit* hits, const float4* pendD, const float4* pendS )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t i  = size_t( y ) * W + x;
    const float4 dd = pendD[ i ];
    if( dd.w < 0.0f )
    {
        Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_DIRECT )[ i ]   = make_float4( 0.f, 0.f, 0.f, 0.f );
        Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_SPECULAR )[ i ] = make_float4( 0.f, 0.f, 0.f, 0.f );
        Plane< float2 >( P, RG_EXT_PLANE_GRADIENT_INPUTS )[ i ]     = make_float2( 0.f, 0.f );
        return;
    }
    const float4 ss   = pendS[ i ];
    v3           diff = mk3( dd.x, dd.y, dd.z ), spec = mk3( ss.x, ss.y, ss.z );
    const bool   shadowTested = rays[ i ].tmax >= rays[ i ].tmin;
    if( shadowTested && hits[ i ].primitiveId != RG_EXT_MISS )
    {
        diff = spec = mk3( 0, 0, 0 );
    }
    const float2 mr     = Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ i ];
    int          rx, ry;
    cbToRegular( x, y, W, &rx, &ry );
    const float4 alb    = Plane< float4 >( P, RG_EXT_PLANE_ALBEDO )[ size_t( ry ) * W + rx ];
    const v3     sc     = SpecularColor( mk3( alb.x, alb.y, alb.z ), mr.x );
    const v3     demod  = mk3( spec.x / fmaxf( 0.01f, sc.x ), spec.y / fmaxf( 0.01f, sc.y ), spec.z / fmaxf( 0.01f, sc.z ) );
    float4&      vd     = Plane< float4 >( P, RG_EXT_PLANE_VIEW_DIRECTION )[ i ];
    Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_DIRECT )[ i ]   = make_float4( diff.x, diff.y, diff.z, 0.f );
    Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_SPECULAR )[ i ] = make_float4( demod.x, demod.y, demod.z, 0.f );
    vd.w                                                        = dd.w;
    Plane< float2 >( P, RG_EXT_PLANE_GRADIENT_INPUTS )[ i ]     = make_float2( luminance( diff ), luminance( spec ) );
}

// Pending state of the indirect pass, 6 float4 per pixel:
//  0: bounce-1 direction, 1/pdf   1: hit position, flags (1 = sky)
//  2: hit normal, directional?    3: hit albedo, -
//  4: unshadowed direct at the hit + emission, -   5: bounce-2 direction, 1/pdf
struct IndirPend
{
    float4* a;
    size_t  n;
    __device__ float4& at( int k, size_t i ) const { return a[ size_t( k ) * n + i ]; }
};

// RT_RAYGEN_INDIRECT_INIT: specular (GGX-VNDF) bounce from the primary surface.
__global__ __launch_bounds__( THREADS ) void k_indirect_a( RgExtFrameParams p, FbPtrs P, LightView lv, RgExtRay* rays, uint8_t* masks, IndirPend pend )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t i = size_t( y ) * W + x;
    Surface      s = FetchSurface( P, W, x, y );
    if( s.isSky )
    {
        rays[ i ]        = NullRay();
        pend.at( 0, i ) = make_float4( 0.f, 0.f, 0.f, -1.0f );
        return;
    }
    s.position          = s.position + s.toViewerDir * 0.01f; // RAY_ORIGIN_LEAK_BIAS
    const uint32_t seed = randomSeed( x, y, p.frameId );
    float          bn[ 4 ];
    BlueNoise8( lv.blueNoise, seed, SALT_SPEC_BOUNCE( 1 ), bn );
    float    oneOverPdf;
    const v3 d = sampleSmithGGX( s.normal, s.toViewerDir, s.roughness, bn[ 0 ], bn[ 1 ], &oneOverPdf, true );
    rays[ i ]  = MakeRay( s.position + s.normal * 0.01f, d, 0.001f, p.rayLength );
    masks[ i ] = uint8_t( IndirectCullMask( p, s.instCustomIndex ) );
    pend.at( 0, i ) = make_float4( d.x, d.y, d.z, oneOverPdf );
}

// Direct light at a bounce hit (processDirectIllumination for
// LIGHT_SAMPLE_METHOD_INDIR, RaygenCommon.h:701-741): RIS without
// visibility; the shadow ray is traced if bounce < maxBounceShadows.
__device__ void IndirectDirect( const RgExtFrameParams& p, const LightView& lv, const Surface& hs, uint32_t seed, int bounce,
                                v3* diffOut, bool* isDirectional, bool* needShadow, v3* lightPos, uint32_t* lightIndex = nullptr )
{
    *diffOut       = mk3( 0, 0, 0 );
    *needShadow    = false;
    *isDirectional = false;
    if( !DirectValid( p, bounce ) )
    {
        return;
    }
    const LightPoint lp = LightPointRnd( seed );
    uint32_t  salt = SALT_LIGHT_CHOOSE_INDIRECT;
    Reservoir regular, dir;
    InitialCandidates( p, lv, hs, seed, salt, lp, true, regular, dir );
    const Reservoir r = CombineInitial( regular, dir, seed, salt );
    if( !isValid( r ) )
    {
        return;
    }
    const LightSample ls = sampleLight( GetLight( lv, r.selected ), hs.position, lp, p.polyLightSpotlightFactor );
    v3                d, sp;
    shadeLight( hs, ls, selectedSampleWeight( r ), true, &d, &sp );
    if( luminance( d + sp ) <= 0.0f )
    {
        return;
    }
    *diffOut       = d;
    *isDirectional = r.selected == 0u;
    *needShadow    = uint32_t( bounce ) < p.maxBounceShadowsLights;
    *lightPos      = ls.position;
    if( lightIndex )
    {
        *lightIndex = r.selected;
    }
}

// Bounce-1 hit: direct light at the hit (shadow ray) and the second bounce.
template< bool TEX >
__global__ __launch_bounds__( THREADS ) void k_indirect_b( RgExtFrameParams p, FbPtrs P, SceneView sc, LightView lv, const RgExtRay* rays1, const RgExtHit* hits1,
                                                           RgExtRay* shadowRays, uint8_t* shadowMasks, RgExtRay* bounceRays, uint8_t* bounceMasks,
                                                           IndirPend pend )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t i = size_t( y ) * W + x;
    shadowRays[ i ] = NullRay();
    bounceRays[ i ] = NullRay();
    if( pend.at( 0, i ).w < 0.0f )
    {
        return;
    }
    const RgExtRay r   = rays1[ i ];
    const RgExtHit hit = hits1[ i ];
    const v3       dir = mk3( r.direction[ 0 ], r.direction[ 1 ], r.direction[ 2 ] );
    const v3       org = mk3( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] );
    if( HitIsSky( p, sc, hit ) )
    {
        pend.at( 1, i ) = make_float4( 0.f, 0.f, 0.f, 1.0f ); // sky
        return;
    }
    TexQuery bq;
    bq.mode             = TQ_BOUNCE; // FIRST_BOUNCE_MIP_BIAS 0 (RtRaygenIndirect.inl:63)
    bq.a                = Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS )[ i ].y;
    const HitInfo  h    = GetHit< TEX >( sc, p, hit.primitiveId, hit.u, hit.v, org, bq );
    const Surface  hs   = HitToSurface( h, dir, CustomIndexOf( sc.geoms[ h.geom ].flags, p.skyGeometry != 0u ) );
    const uint32_t seed = randomSeed( x, y, p.frameId );
    v3             diff, lightPos = mk3( 0, 0, 0 );
    bool           isDir, needShadow;
    uint32_t       lightIndex = LIGHT_INDEX_NONE;
    IndirectDirect( p, lv, hs, seed, 1, &diff, &isDir, &needShadow, &lightPos, &lightIndex );
    if( needShadow )
    {
        shadowRays[ i ]  = ShadowRay( hs, lightPos );
        shadowMasks[ i ] = uint8_t( ShadowCullMask( p, hs.instCustomIndex, lightIndex ) );
    }
    const v3 emis = h.emission * p.emissionMapBoost;
    // second diffuse bounce (getDiffuseBounce, RtRaygenIndirect.inl:57-62)
    float bn[ 4 ];
    BlueNoise8( lv.blueNoise, seed, SALT_DIFF_BOUNCE( 2 ), bn );
    float    pdf2;
    const v3 d2     = sampleOrientedHemisphere( hs.normal, bn[ 0 ], bn[ 1 ], &pdf2 );
    bounceRays[ i ]  = MakeRay( hs.position + hs.normal * 0.01f, d2, 0.001f, p.rayLength );
    bounceMasks[ i ] = uint8_t( IndirectCullMask( p, hs.instCustomIndex ) );
    pend.at( 1, i ) = make_float4( h.position.x, h.position.y, h.position.z, 0.0f );
    pend.at( 2, i ) = make_float4( h.normal.x, h.normal.y, h.normal.z, isDir ? 1.0f : 0.0f );
    pend.at( 3, i ) = make_float4( h.albedo.x, h.albedo.y, h.albedo.z, needShadow ? 1.0f : 0.0f );
    pend.at( 4, i ) = make_float4( diff.x, diff.y, diff.z, 0.0f );
    pend.at( 5, i ) = make_float4( d2.x, d2.y, d2.z, pdf2 );
    ( void )emis;
}

// Second-bounce hit and the final indirect sample; shading of the primary
// surface with it (RtRaygenIndirect.inl:66-161, :402-433 without reuse).
// The GI_REUSE record of pixel i: what the spatial reuse of other pixels
// reads of it (its G-buffer surface and its initial sample, normals decoded
// once here instead of at every reuse), 64 bytes in one cache line.
__device__ __forceinline__ void StoreGiReuse( const FbPtrs& P, size_t i, uint4 sA, uint2 sB )
{
    const float4 x1  = Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION )[ i ];
    const float  d   = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ i ];
    const v3     n   = DecodeNormal( Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL )[ i ] );
    const v3     sn  = DecodeNormal( sA.w );
    auto         u   = []( float f ) { return __float_as_uint( f ); };
    uint4*       rec = Plane< uint4 >( P, RG_EXT_PLANE_GI_REUSE ) + 4 * i;
    rec[ 0 ]         = make_uint4( u( x1.x ), u( x1.y ), u( x1.z ), u( d ) );
    rec[ 1 ]         = make_uint4( u( n.x ), u( n.y ), u( n.z ), Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ i ] );
    rec[ 2 ]         = make_uint4( sA.x, sA.y, sA.z, sB.y );
    rec[ 3 ]         = make_uint4( u( sn.x ), u( sn.y ), u( sn.z ), sB.x );
}

template< bool TEX >
__global__ __launch_bounds__( THREADS ) void k_indirect_c( RgExtFrameParams p, FbPtrs P, SceneView sc, LightView lv, const RgExtHit* shadowHits,
                                                           const RgExtRay* bounceRays, const RgExtHit* bounceHits, const RgExtHit* hits1, IndirPend pend )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t i  = size_t( y ) * W + x;
    const float4 b1 = pend.at( 0, i );
    if( b1.w < 0.0f )
    {
        // sky pixel: emptySampleIndirect with 1/pdf 0 (RtRaygenIndirect.inl:225-229)
        Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_INDIR )[ i ]        = make_float4( 0.f, 0.f, 0.f, 0.f );
        Plane< uint4 >( P, RG_EXT_PLANE_INDIR_INITIAL_SAMPLE )[ i ]     = make_uint4( 0u, 0u, 0u, 0u );
        Plane< uint2 >( P, RG_EXT_PLANE_INDIR_INITIAL_RADIANCE )[ i ] = make_uint2( 0u, 0u );
        StoreGiReuse( P, i, make_uint4( 0u, 0u, 0u, 0u ), make_uint2( 0u, 0u ) );
        return;
    }
    Surface s  = FetchSurface( P, W, x, y );
    s.position = s.position + s.toViewerDir * 0.01f;
    const v3 dir1 = mk3( b1.x, b1.y, b1.z );
    v3       samplePos, sampleRad, sampleNrm;
    const float4 hp = pend.at( 1, i );
    if( hp.w > 0.5f )
    {
        // sky sample (RtRaygenIndirect.inl:123-131)
        samplePos = s.position + dir1 * MAX_RAY_LEN;
        sampleNrm = -dir1;
        sampleRad = SkyColor( p );
    }
    else
    {
        const float4 hn = pend.at( 2, i ), ha = pend.at( 3, i ), hd = pend.at( 4, i ), b2 = pend.at( 5, i );
        v3           diffuse = mk3( hd.x, hd.y, hd.z );
        if( ha.w > 0.5f && shadowHits[ i ].primitiveId != RG_EXT_MISS )
        {
            diffuse = mk3( 0, 0, 0 );
        }
        if( hn.w < 0.5f )
        {
            // SHIPPING_HACK: non-directional light at a bounce is clamped to 20
            diffuse = mk3( clampf( diffuse.x, 0.0f, 20.0f ), clampf( diffuse.y, 0.0f, 20.0f ), clampf( diffuse.z, 0.0f, 20.0f ) );
        }
        // emission of the bounce-1 hit
        const RgExtHit h1   = hits1[ i ];
        TexQuery       q1;
        q1.mode             = TQ_BOUNCE;
        q1.a                = s.roughness;
        const HitInfo  hi1  = GetHit< TEX >( sc, p, h1.primitiveId, h1.u, h1.v, s.position + s.normal * 0.01f, q1 );
        const v3       emis = hi1.emission * p.emissionMapBoost;
        // processSecondDiffuseBounce (RtRaygenIndirect.inl:82-101)
        const v3       d2 = mk3( b2.x, b2.y, b2.z );
        const RgExtHit h2 = bounceHits[ i ];
        v3             second;
        if( HitIsSky( p, sc, h2 ) )
        {
            second = SkyColor( p ) * b2.w;
        }
        else
        {
            const RgExtRay br   = bounceRays[ i ];
            TexQuery       q2;
            q2.mode             = TQ_BOUNCE;
            q2.a                = hi1.roughness;
            q2.b                = 32.0f; // SECOND_BOUNCE_MIP_BIAS (RtRaygenIndirect.inl:64)
            const HitInfo  hi2  = GetHit< TEX >( sc, p, h2.primitiveId, h2.u, h2.v, mk3( br.origin[ 0 ], br.origin[ 1 ], br.origin[ 2 ] ), q2 );
            const Surface  hs2  = HitToSurface( hi2, d2, CustomIndexOf( sc.geoms[ hi2.geom ].flags, p.skyGeometry != 0u ) );
            const uint32_t seed = randomSeed( x, y, p.frameId );
            v3             diff2, lp;
            bool           isDir2, needShadow2;
            // direct light at the second bounce, unshadowed (a shadow ray is
            // only requested there when maxBounceShadows > 2)
            IndirectDirect( p, lv, hs2, seed, 2, &diff2, &isDir2, &needShadow2, &lp );
            if( !isDir2 )
            {
                diff2 = mk3( clampf( diff2.x, 0.0f, 20.0f ), clampf( diff2.y, 0.0f, 20.0f ), clampf( diff2.z, 0.0f, 20.0f ) );
            }
            second = ( hi2.emission * p.emissionMapBoost + diff2 ) * hi2.albedo * b2.w;
        }
        diffuse   = diffuse + second;
        samplePos = mk3( hp.x, hp.y, hp.z );
        sampleNrm = mk3( hn.x, hn.y, hn.z );
        sampleRad = ( emis + diffuse ) * mk3( ha.x, ha.y, ha.z );
    }
    // restirIndirect_StoreInitialSample (ReservoirIndirect.h:160-177)
    const uint4 sA = make_uint4( __float_as_uint( samplePos.x ), __float_as_uint( samplePos.y ), __float_as_uint( samplePos.z ), EncodeNormal( sampleNrm ) );
    const uint2 sB = make_uint2( EncodeE5B9G9R9( sampleRad ), __float_as_uint( b1.w ) );
    Plane< uint4 >( P, RG_EXT_PLANE_INDIR_INITIAL_SAMPLE )[ i ]   = sA;
    Plane< uint2 >( P, RG_EXT_PLANE_INDIR_INITIAL_RADIANCE )[ i ] = sB;
    StoreGiReuse( P, i, sA, sB );
}

// ------------------------------------------------------------ ReSTIR GI

// SampleIndirect / ReservoirIndirect (ReservoirIndirect.h:23-36)
struct SampleIndir
{
    v3 position, normal, radiance;
};

struct ReservoirIndir
{
    SampleIndir sel;
    float       selTargetPdf;
    float       weightSum;
    uint32_t    M;
};

__device__ __forceinline__ ReservoirIndir EmptyReservoirIndir()
{
    ReservoirIndir r;
    r.sel.position = r.sel.normal = r.sel.radiance = mk3( 0, 0, 0 );
    r.selTargetPdf = r.weightSum = 0.0f;
    r.M                          = 0u;
    return r;
}

// calcSelectedSampleWeightIndirect (:57-60)
__device__ __forceinline__ float SelectedWeightIndir( const ReservoirIndir& r )
{
    return safePositiveRcp( r.selTargetPdf ) * ( r.weightSum / float( r.M > 1u ? r.M : 1u ) );
}

// normalizeReservoirIndirect (:62-68)
__device__ __forceinline__ void NormalizeIndir( ReservoirIndir& r, uint32_t maxM )
{
    r.weightSum /= float( r.M > 1u ? r.M : 1u );
    r.M = r.M < maxM ? r.M : maxM;
    r.weightSum *= float( r.M );
}

// updateCombinedReservoirIndirect (:90-104)
__device__ __forceinline__ void CombineIndir( ReservoirIndir& c, const ReservoirIndir& b, float rnd )
{
    const float w = b.weightSum;
    c.weightSum += w;
    c.M += b.M;
    if( rnd * c.weightSum < w )
    {
        c.sel          = b.sel;
        c.selTargetPdf = b.selTargetPdf;
    }
}

// updateCombinedReservoirIndirect_newSurf (:106-119)
__device__ __forceinline__ void CombineIndirNewSurf( ReservoirIndir& c, const ReservoirIndir& b, float tpB, float rnd )
{
    const float w = tpB * safePositiveRcp( b.selTargetPdf ) * b.weightSum;
    c.weightSum += w;
    c.M += b.M;
    if( rnd * c.weightSum < w )
    {
        c.sel          = b.sel;
        c.selTargetPdf = tpB;
    }
}

// loadInitialSampleAsReservoir (RtRaygenIndirect.inl:246-255):
// updateReservoirIndirect (ReservoirIndirect.h:70-84) with rnd = 0.5
__device__ __forceinline__ ReservoirIndir InitialAsReservoir( v3 position, v3 normal, uint32_t radianceCode, float oneOverSourcePdf )
{
    SampleIndir    s{ position, normal, DecodeE5B9G9R9( radianceCode ) };
    const float    tp               = luminance( s.radiance ); // targetPdfForIndirectSample
    ReservoirIndir r                = EmptyReservoirIndir();
    const float    w                = tp * oneOverSourcePdf;
    r.weightSum += w;
    r.M += 1u;
    if( 0.5f * r.weightSum < w )
    {
        r.sel          = s;
        r.selTargetPdf = tp;
    }
    return r;
}

__device__ __forceinline__ ReservoirIndir InitialAsReservoir( uint4 a, uint2 b )
{
    return InitialAsReservoir( mk3( __uint_as_float( a.x ), __uint_as_float( a.y ), __uint_as_float( a.z ) ), DecodeNormal( a.w ), b.x,
                               __uint_as_float( b.y ) );
}

__device__ __forceinline__ ReservoirIndir InitialAsReservoir( const FbPtrs& P, size_t i )
{
    return InitialAsReservoir( Plane< uint4 >( P, RG_EXT_PLANE_INDIR_INITIAL_SAMPLE )[ i ], Plane< uint2 >( P, RG_EXT_PLANE_INDIR_INITIAL_RADIANCE )[ i ] );
}

__device__ __forceinline__ ReservoirIndir LoadReservoirIndir( const FbPtrs& P, size_t i, bool prev )
{
    const uint4    a = Plane< uint4 >( P, RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, prev )[ i ];
    const uint4    b = Plane< uint4 >( P, RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, prev )[ i ];
    ReservoirIndir r;
    r.sel.position = mk3( __uint_as_float( a.x ), __uint_as_float( a.y ), __uint_as_float( a.z ) );
    r.sel.normal   = DecodeNormal( a.w );
    r.sel.radiance = DecodeE5B9G9R9( b.x );
    r.selTargetPdf = __uint_as_float( b.y );
    r.weightSum    = __uint_as_float( b.z );
    r.M            = b.w;
    return r;
}

// restirIndirect_StoreReservoir (:179-212): non-finite or negative weights store zeros
__device__ __forceinline__ void StoreReservoirIndir( const FbPtrs& P, size_t i, const ReservoirIndir& r )
{
    uint4 a = make_uint4( 0u, 0u, 0u, 0u ), b = make_uint4( 0u, 0u, 0u, 0u );
    if( isfinite( r.weightSum ) && r.weightSum >= 0.0f )
    {
        a = make_uint4( __float_as_uint( r.sel.position.x ), __float_as_uint( r.sel.position.y ), __float_as_uint( r.sel.position.z ), EncodeNormal( r.sel.normal ) );
        b = make_uint4( EncodeE5B9G9R9( r.sel.radiance ), __float_as_uint( r.selTargetPdf ), __float_as_uint( r.weightSum ), r.M );
    }
    Plane< uint4 >( P, RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE )[ i ] = a;
    Plane< uint4 >( P, RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT )[ i ] = b;
}

// getDiffuseWeight (RtRaygenIndirect.inl:207-212): smoothstep(MIN_GGX_ROUGHNESS, 0.5 + 0.25, roughness)
__device__ __forceinline__ float DiffuseWeight( float roughness )
{
    const float t = clampf( ( roughness - 0.005f ) / ( 0.75f - 0.005f ), 0.0f, 1.0f );
    return t * t * ( 3.0f - 2.0f * t );
}

// testSurfaceForReuseIndirect (RtRaygenIndirect.inl:189-203)
__device__ __forceinline__ bool ReuseIndirOk( int lo, int hi, int H, int qx, int qy, float dCur, float dOther, v3 nCur, v3 nOther )
{
    return inRenderArea( qx, qy, lo, hi, H ) && ( fabsf( dCur - dOther ) / fabsf( dCur ) < 0.05f ) && ( dot( nCur, nOther ) > 0.0f );
}

// getGeometryFactorClamped (Light.h:135-138) with calcDirectionAndLengthSafe (Utils.h:93-101)
__device__ __forceinline__ float GeometryFactorSafe( v3 n, v3 from, v3 to )
{
    v3          d   = to - from;
    const float len = fmaxf( length( d ), 0.001f );
    d               = d / len;
    return fmaxf( 0.0f, dot( n, d ) ) * safePositiveRcp( square( len ) );
}

// RT_RAYGEN_INDIRECT_FINAL (RtRaygenIndirect.inl:257-434): temporal and
// spatial reuse of the indirect samples, then shading of the primary
// surface with the selected sample.
__global__ __launch_bounds__( THREADS ) void k_indirect_final( RgExtFrameParams p, FbPtrs P )
{
    const int x = blockIdx.x * BX + threadIdx.x % BX, y = blockIdx.y * BY + threadIdx.x / BX;
    const int W = p.renderWidth, H = p.renderHeight;
    if( x >= W || y >= H )
    {
        return;
    }
    const size_t i = size_t( y ) * W + x;
    Surface      s = FetchSurface( P, W, x, y );
    if( s.isSky )
    {
        StoreReservoirIndir( P, i, EmptyReservoirIndir() );
        return;
    }
    s.position          = s.position + s.toViewerDir * 0.01f; // RAY_ORIGIN_LEAK_BIAS
    const uint32_t seed = randomSeed( x, y, p.frameId );
    uint32_t       salt = SALT_RESAMPLE_INDIRECT_BASE;

    ReservoirIndir comb = InitialAsReservoir( P, i );

    int lo, hi;
    cbRenderArea( x, W, &lo, &hi );
    const float4 mot      = Plane< float4 >( P, RG_EXT_PLANE_MOTION )[ i ];
    const float  depthCur = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD )[ i ];
    // getPrevScreenPos (ShaderCommonGLSLFunc.h:271-277) over the checkerboard half width
    const float  px       = ( ( float( x ) + 0.5f ) * ( 1.0f / ( float( W ) / 2.0f ) ) + mot.x ) * ( float( W ) / 2.0f );
    const float  py       = ( ( float( y ) + 0.5f ) * ( 1.0f / float( H ) ) + mot.y ) * float( H );
    const float  dw       = DiffuseWeight( s.roughness );
    int          spatialCount = int( 6.0f * dw ); // SPATIAL_SAMPLES_INDIR

    // TEMPORAL_SAMPLES_INDIR = 1, TEMPORAL_RADIUS_INDIR_MAX = 2
    {
        float r4[ 4 ];
        rnd8_4( seed, salt++, r4 );
        const float ox = ( r4[ 0 ] * 2.0f - 1.0f ) * square( dw );
        const float oy = ( r4[ 1 ] * 2.0f - 1.0f ) * square( dw );
        const int   qx = int( floorf( px + ox * 2.0f ) );
        const int   qy = int( floorf( py + oy * 2.0f ) );
        // the previous frame's planes in one round trip, at the pixel clamped
        // into the image: a pixel outside fails the render-area check
        const size_t   qi    = size_t( min( max( qy, 0 ), H - 1 ) ) * W + min( max( qx, 0 ), W - 1 );
        const uint32_t skyQ  = Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY )[ qi ];
        const float    dPrev = Plane< float >( P, RG_EXT_PLANE_DEPTH_WORLD, true )[ qi ];
        const uint32_t nPrev = Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL, true )[ qi ];
        // antilag of the indirect signal in the gradient history (strata pixels)
        const int      GW    = ( W + 1 ) / 3 > 1 ? ( W + 1 ) / 3 : 1;
        const int      GH    = ( H + 1 ) / 3 > 1 ? ( H + 1 ) / 3 : 1;
        const float4   gh    = LdPlane( Plane< float4 >( P, RG_EXT_PLANE_DIS_GRADIENT_HISTORY ), GW, GH, qx / 3, qy / 3, make_float4( 0.f, 0.f, 0.f, 0.f ) );
        ReservoirIndir t     = LoadReservoirIndir( P, qi, true );
        bool           use   = skyQ == 0u && ReuseIndirOk( lo, hi, H, qx, qy, depthCur, dPrev - mot.z, s.normal, DecodeNormal( nPrev ) );
        if( use && gh.y > 0.25f )
        {
            spatialCount++;
            use = false;
        }
        if( use )
        {
            NormalizeIndir( t, 20u );
            CombineIndir( comb, t, rnd16( seed, salt++ ) );
        }
    }

    // spatial reuse with the Jacobian of the sample's solid angle
    {
        const float radius  = mixf( 8.0f, 32.0f, clampf( float( H ) / 1080.0f, 0.0f, 1.0f ) );
        uint32_t    nobiasM = comb.M;
        for( int k = 0; k < spatialCount; k++ )
        {
            float r4[ 4 ];
            rnd8_4( seed, salt++, r4 );
            const int qx = x + int( ( r4[ 0 ] * 2.0f - 1.0f ) * radius );
            const int qy = y + int( ( r4[ 1 ] * 2.0f - 1.0f ) * radius );
            // the neighbour's GI_REUSE record (one cache line), at the pixel
            // clamped into the image; a pixel outside fails the reuse test
            // (its render-area check), so the clamped values are never used
            const size_t  qi  = size_t( min( max( qy, 0 ), H - 1 ) ) * W + min( max( qx, 0 ), W - 1 );
            const uint4*  rec = Plane< uint4 >( P, RG_EXT_PLANE_GI_REUSE ) + 4 * qi;
            const uint4   r0 = rec[ 0 ], r1 = rec[ 1 ], r2 = rec[ 2 ], r3 = rec[ 3 ];
            auto          f  = []( uint32_t v ) { return __uint_as_float( v ); };
            const float4  x1q = make_float4( f( r0.x ), f( r0.y ), f( r0.z ), 0.0f );
            if( r1.w != 0u || !ReuseIndirOk( lo, hi, H, qx, qy, depthCur, f( r0.w ), s.normal, mk3( f( r1.x ), f( r1.y ), f( r1.z ) ) ) )
            {
                continue;
            }
            const ReservoirIndir rq = InitialAsReservoir( mk3( f( r2.x ), f( r2.y ), f( r2.z ) ), mk3( f( r3.x ), f( r3.y ), f( r3.z ) ), r3.w, f( r2.w ) );
            float oneOverJacobian    = safePositiveRcp( GeometryFactorSafe( rq.sel.normal, rq.sel.position, s.position ) ) *
                                    GeometryFactorSafe( rq.sel.normal, rq.sel.position, mk3( x1q.x, x1q.y, x1q.z ) );
            oneOverJacobian          = clampf( oneOverJacobian, 0.0f, 1.0f ); // SHIPPING_HACK
            const float tpCur        = luminance( rq.sel.radiance ) * oneOverJacobian;
            CombineIndirNewSurf( comb, rq, tpCur, rnd16( seed, salt++ ) );
            if( tpCur > 0.0f )
            {
                nobiasM += rq.M;
            }
        }
        comb.M = nobiasM;
    }
    StoreReservoirIndir( P, i, comb );

    // shade (RtRaygenIndirect.inl:164-182)
    const float oneOverPdf = SelectedWeightIndir( comb );
    v3          diff = mk3( 0, 0, 0 ), spec = mk3( 0, 0, 0 );
    const v3    l    = safeNormalize( comb.sel.position - s.position );
    const float nl   = dot( s.normal, l );
    if( nl > 0.0f )
    {
        diff = comb.sel.radiance * ( nl * ( 1.0f / RG_PI ) ) * oneOverPdf;
        spec = ( comb.sel.radiance * nl ) * evalBRDFSmithGGX( s.normal, s.toViewerDir, l, s.roughness, s.specularColor, true ) * oneOverPdf;
    }
    float4&  us     = Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_SPECULAR )[ i ];
    const v3 direct = mk3( us.x, us.y, us.z );
    if( luminance( direct ) < luminance( spec ) )
    {
        Plane< float4 >( P, RG_EXT_PLANE_VIEW_DIRECTION )[ i ].w = length( comb.sel.position - s.position );
    }
    const v3 sc2 = s.specularColor;
    us           = make_float4( direct.x + spec.x / fmaxf( 0.01f, sc2.x ), direct.y + spec.y / fmaxf( 0.01f, sc2.y ), direct.z + spec.z / fmaxf( 0.01f, sc2.z ), 0.0f );
    Plane< float4 >( P, RG_EXT_PLANE_UNFILTERED_INDIR )[ i ] = make_float4( diff.x, diff.y, diff.z, 0.0f );
}


// ------------------------------------------------------------ gradient samples

// getAntilagAlpha (ShaderCommonGLSLFunc.h:383-390)
__device__ __forceinline__ float AntilagAlpha( float gradSample, float normFactor )
{
    const float lambda = normFactor > 0.0001f ? clampf( fabsf( gradSample ) / normFactor, 0.0f, 1.0f ) : 0.0f;
    return clampf( lambda, 0.0f, 1.0f );
}

struct GradPend
{
    float4* lum; // forward-projected diffuse / specular luminance before visibility, previous luminance
    uint4*  pix; // prev shading pixel x, y, strata position, state (0 rejected, 1 no shadow ray, 2 shadow ray)
};

__device__ __forceinline__ void StoreGradient( const FbPtrs& P, int GW, int gx, int gy, float4 g, uint32_t strataPos )
{
    Plane< float4 >( P, RG_EXT_PLANE_DIS_PING_GRADIENT )[ size_t( gy ) * GW + gx ] = g;
    Plane< uint32_t >( P, RG_EXT_PLANE_GRADIENT_PREV_PIX )[ size_t( gy ) * GW + gx ] = strataPos;
}

// RtGradients.rgen:44-131: per 3x3 stratum, the brightest pixel of the
// previous frame (not the one used last frame) is re-shaded with its old
// reservoir, light point and seed on the surface as it is now.
__global__ __launch_bounds__( THREADS ) void k_grad_a( RgExtFrameParams p, FbPtrs P, SceneView sc, LightView lv, const uint32_t* geomPrevToCur,
                                                       uint32_t geomPrevCount, uint32_t geomCurCount, RgExtRay* rays, uint8_t* masks, GradPend pend )
{
    const int W = p.renderWidth, H = p.renderHeight;
    const int GW = ( W + 1 ) / 3 > 1 ? ( W + 1 ) / 3 : 1, GH = ( H + 1 ) / 3 > 1 ? ( H + 1 ) / 3 : 1;
    const int gx = blockIdx.x * BX + threadIdx.x % BX, gy = blockIdx.y * BY + threadIdx.x / BX;
    if( gx >= GW || gy >= GH )
    {
        return;
    }
    const size_t gi = size_t( gy ) * GW + gx;
    rays[ gi ]      = NullRay();
    pend.pix[ gi ]  = make_uint4( 0u, 0u, 10u, 0u );
    const float4 zero = make_float4( 0.f, 0.f, 0.f, 0.f );
    if( gx * 3 >= W || gy * 3 >= H )
    {
        return;
    }
    if( p.gradientMultDiffuse < 0.001f && p.gradientMultIndirect < 0.001f && p.gradientMultSpecular < 0.001f )
    {
        StoreGradient( P, GW, gx, gy, zero, 10u );
        return;
    }
    const uint32_t prevStrataPos = Plane< uint32_t >( P, RG_EXT_PLANE_GRADIENT_PREV_PIX )[ gi ];
    int            psx = 0, psy = 0, csx = 0, csy = 0;
    uint32_t       curStrataPos = 0;
    float2         prevLum      = make_float2( 0.f, 0.f );
    for( int yy = 0; yy < 3; yy++ )
    {
        for( int xx = 0; xx < 3; xx++ )
        {
            if( uint32_t( yy * 3 + xx ) == prevStrataPos )
            {
                continue;
            }
            const int    cx  = gx * 3 + xx, cy = gy * 3 + yy;
            const float4 mot = LdPlane( Plane< float4 >( P, RG_EXT_PLANE_MOTION ), W, H, cx, cy, zero );
            const int    qx  = int( floorf( ( ( float( cx ) + 0.5f ) * ( 1.0f / ( float( W ) / 2.0f ) ) + mot.x ) * ( float( W ) / 2.0f ) ) );
            const int    qy  = int( floorf( ( ( float( cy ) + 0.5f ) * ( 1.0f / float( H ) ) + mot.y ) * float( H ) ) );
            const float2 l   = LdPlane( Plane< float2 >( P, RG_EXT_PLANE_GRADIENT_INPUTS, true ), W, H, qx, qy, make_float2( 0.f, 0.f ) );
            if( fmaxf( l.x, l.y ) > fmaxf( prevLum.x, prevLum.y ) )
            {
                psx = qx, psy = qy, csx = cx, csy = cy;
                curStrataPos = uint32_t( yy * 3 + xx );
                prevLum      = l;
            }
        }
    }
    int lo, hi;
    cbRenderArea( csx, W, &lo, &hi );
    if( !inRenderArea( psx, psy, lo, hi, H ) || LdPlane( Plane< uint32_t >( P, RG_EXT_PLANE_IS_SKY ), W, H, csx, csy, 0u ) != 0u ||
        ( prevLum.x <= 0.0f && prevLum.y <= 0.0f ) )
    {
        StoreGradient( P, GW, gx, gy, zero, 10u );
        return;
    }
    const size_t   pi      = size_t( psy ) * W + psx;
    const uint32_t oldSeed = randomSeed( psx, psy, p.frameId - 1u );

    // fetchGbufferSurface_NoAlbedoViewDir_Prev (Surface.inl:72-90)
    Surface ps;
    ps.isSky           = false;
    ps.albedo          = mk3( 1, 1, 1 );
    const float4 pos   = Plane< float4 >( P, RG_EXT_PLANE_SURFACE_POSITION, true )[ pi ];
    const float2 mr    = Plane< float2 >( P, RG_EXT_PLANE_METALLIC_ROUGHNESS, true )[ pi ];
    ps.position        = mk3( pos.x, pos.y, pos.z );
    ps.instCustomIndex = __float_as_uint( pos.w );
    ps.specularColor   = SpecularColor( ps.albedo, mr.x );
    ps.roughness       = mr.y;
    ps.normal          = DecodeNormal( Plane< uint32_t >( P, RG_EXT_PLANE_NORMAL, true )[ pi ] );

    // unpackPrevVisibilityBuffer (VertexData.inl:636-658): the previous hit's
    // triangle as it is in this frame
    const uint4 vb      = Plane< uint4 >( P, RG_EXT_PLANE_VISIBILITY_BUFFER, true )[ pi ];
    bool        matched = false;
    if( vb.x == 1u )
    {
        const uint32_t pg    = vb.y & 0xFFFu;
        const uint32_t local = vb.y >> 12;
        const uint32_t cg    = pg < geomPrevCount ? geomPrevToCur[ pg ] : 0xFFFFFFFFu;
        if( cg < geomCurCount && local < sc.geoms[ cg ].triCount )
        {
            const float* t  = sc.tris + size_t( sc.geoms[ cg ].firstTri + local ) * 9;
            const float  u  = __uint_as_float( vb.z ), v = __uint_as_float( vb.w );
            const float  b0 = 1.0f - u - v;
            ps.position     = mk3( t[ 0 ], t[ 1 ], t[ 2 ] ) * b0 + mk3( t[ 3 ], t[ 4 ], t[ 5 ] ) * u + mk3( t[ 6 ], t[ 7 ], t[ 8 ] ) * v;
            matched         = true;
        }
    }
    ps.toViewerDir = normalize( mk3( p.cameraPositionPrev[ 0 ], p.cameraPositionPrev[ 1 ], p.cameraPositionPrev[ 2 ] ) - ps.position );

    float lumD = 0.0f, lumS = 0.0f;
    uint32_t state = 1u;
    if( matched )
    {
        Reservoir r = unpackReservoir( Plane< uint4 >( P, RG_EXT_PLANE_RESERVOIRS, true )[ pi ] );
        if( isValid( r ) )
        {
            const uint32_t cur = r.selected < lv.prevCount ? lv.prevToCur[ r.selected ] : 0xFFFFFFFFu;
            if( cur != 0xFFFFFFFFu && cur != LIGHT_INDEX_NONE && DirectValid( p, 0 ) )
            {
                r.selected = cur;
                // processDirectIllumination for LIGHT_SAMPLE_METHOD_GRADIENTS
                // (RaygenCommon.h:742-760) -> traceDirectIllumination (:639-674)
                const LightPoint lp = LightPointRnd( oldSeed );
                const LightSample ls = sampleLight( GetLight( lv, cur ), ps.position, lp, p.polyLightSpotlightFactor );
                v3                d, sp;
                shadeLight( ps, ls, selectedSampleWeight( r ), false, &d, &sp );
                if( luminance( d + sp ) > 0.0f )
                {
                    lumD = luminance( d );
                    lumS = luminance( sp );
                    if( 0u < p.maxBounceShadowsLights )
                    {
                        rays[ gi ]  = ShadowRay( ps, ls.position );
                        masks[ gi ] = uint8_t( ShadowCullMask( p, ps.instCustomIndex, cur ) );
                        state       = 2u;
                    }
                }
            }
        }
    }
    pend.lum[ gi ] = make_float4( lumD, lumS, prevLum.x, prevLum.y );
    pend.pix[ gi ] = make_uint4( uint32_t( psx ), uint32_t( psy ), curStrataPos, state );
}

// RtGradients.rgen:133-181: gradients of the three signals, temporally smoothed.
__global__ __launch_bounds__( THREADS ) void k_grad_b( RgExtFrameParams p, FbPtrs P, const RgExtHit* hits, GradPend pend )
{
    const int W = p.renderWidth, H = p.renderHeight;
    const int GW = ( W + 1 ) / 3 > 1 ? ( W + 1 ) / 3 : 1, GH = ( H + 1 ) / 3 > 1 ? ( H + 1 ) / 3 : 1;
    const int gx = blockIdx.x * BX + threadIdx.x % BX, gy = blockIdx.y * BY + threadIdx.x / BX;
    if( gx >= GW || gy >= GH )
    {
        return;
    }
    const size_t gi = size_t( gy ) * GW + gx;
    const uint4  px = pend.pix[ gi ];
    if( px.w == 0u )
    {
        return; // rejected: k_grad_a stored the zero gradient
    }
    const float4 l    = pend.lum[ gi ];
    const float  vis  = ( px.w == 2u && hits[ gi ].primitiveId != RG_EXT_MISS ) ? 0.0f : 1.0f;
    const float  fwdD = l.x * vis, fwdS = l.y * vis;
    float        g[ 3 ];
    g[ 0 ] = AntilagAlpha( fwdD - l.z, fmaxf( fwdD, l.z ) );
    {
        // indirect: only a sudden darkening of the direct diffuse, unless bright
        float       gs         = fmaxf( 0.0f, l.z - fwdD );
        const size_t q         = size_t( px.y ) * W + px.x;
        float4       ia;
        if( P.half )
        {
            const _Float16* h = Plane< _Float16 >( P, RG_EXT_PLANE_INDIR_ACCUM, true ) + q * 4;
            ia                = make_float4( float( h[ 0 ] ), float( h[ 1 ] ), float( h[ 2 ] ), 0.0f );
        }
        else
        {
            ia = Plane< float4 >( P, RG_EXT_PLANE_INDIR_ACCUM, true )[ q ];
        }
        const float brightness = luminance( mk3( ia.x, ia.y, ia.z ) );
        const float t          = clampf( ( brightness - 1.0f ) / ( 10.0f - 1.0f ), 0.0f, 1.0f );
        gs *= 1.0f - t * t * ( 3.0f - 2.0f * t );
        g[ 1 ] = AntilagAlpha( gs, fmaxf( fwdD, l.z ) );
    }
    g[ 2 ] = AntilagAlpha( fwdS - l.w, fmaxf( fwdS, l.w ) );
    g[ 0 ] *= p.gradientMultDiffuse;
    g[ 1 ] *= p.gradientMultIndirect;
    g[ 2 ] *= p.gradientMultSpecular;
    const float4 hp = LdPlane( Plane< float4 >( P, RG_EXT_PLANE_DIS_GRADIENT_HISTORY ), GW, GH, int( px.x ) / 3, int( px.y ) / 3, make_float4( 0.f, 0.f, 0.f, 0.f ) );
    const float4 out = make_float4( mixf( g[ 0 ], hp.x, 0.5f ), mixf( g[ 1 ], hp.y, 0.5f ), mixf( g[ 2 ], hp.z, 0.5f ), 0.0f );
    StoreGradient( P, GW, gx, gy, out, px.z );
}

} // namespace

PathTracer::~PathTracer()
{
    for( auto& e : traceEv )
    {
        if( e )
        {
            ( void )hipEventDestroy( e );
        }
    }
}

float PathTracer::TraversalMs() const
{
    float total = 0.0f;
    for( uint32_t k = 0; k < traceLaunches; k++ )
    {
        float ms = 0.0f;
        if( hipEventElapsedTime( &ms, traceEv[ 2 * k ], traceEv[ 2 * k + 1 ] ) == hipSuccess )
        {
            total += ms;
        }
    }
    return total;
}

void PathTracer::TraversalCounts( uint64_t* nodes, uint64_t* tris, uint32_t* errors ) const
{
    uint64_t v[ 3 ] = {};
    if( traceLaunches > 0 )
    {
        RG_HIP_CHECK( hipMemcpy( v, frameAccum.ptr, sizeof( v ), hipMemcpyDeviceToHost ) );
    }
    *nodes  = collectStats ? v[ 0 ] : 0;
    *tris   = collectStats ? v[ 1 ] : 0;
    *errors = uint32_t( v[ 2 ] );
}

namespace
{
// The seeded noise table is constant: one copy serves every path tracer of
// the process.  It is never freed (8 MB), so no path tracer outlives it.
std::mutex g_noiseMutex;
uint32_t*  g_noise = nullptr;

const uint32_t* SharedBlueNoise( hipStream_t s )
{
    std::lock_guard< std::mutex > lock( g_noiseMutex );
    if( !g_noise )
    {
        void* p = nullptr;
        RG_HIP_CHECK( hipMalloc( &p, size_t( 128u ) * 128u * 128u * 4u ) );
        k_blue_noise<<< DivUp( 128u * 128u * 128u, THREADS ), THREADS, 0, s >>>( static_cast< uint32_t* >( p ) );
        CheckLaunch( "k_blue_noise" );
        RG_HIP_CHECK( hipStreamSynchronize( s ) );
        g_noise = static_cast< uint32_t* >( p );
    }
    return g_noise;
}
} // namespace

void PrimaryRays( const RgExtFrameParams& p, RgExtRay* rays, hipStream_t s )
{
    RG_REQUIRE( p.renderWidth > 0 && p.renderHeight > 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render size must be > 0" );
    const dim3 grid( DivUp( p.renderWidth, BX ), DivUp( p.renderHeight, BY ) );
    k_primary_rays<<< grid, THREADS, 0, s >>>( p, rays );
    CheckLaunch( "k_primary_rays" );
}

size_t Wavefront::BytesFor( size_t n )
{
    return n * ( 2 * sizeof( RgExtRay ) + 2 * sizeof( RgExtHit ) + 16 * 6 + sizeof( RgExtRay ) + sizeof( RgExtHit ) + 3 );
}

void Wavefront::Reserve( size_t n )
{
    if( n <= capPx )
    {
        return;
    }
    rays.Reserve( n * sizeof( RgExtRay ) );
    rays2.Reserve( n * sizeof( RgExtRay ) );
    hits.Reserve( n * sizeof( RgExtHit ) );
    hits2.Reserve( n * sizeof( RgExtHit ) );
    pendA.Reserve( n * 16 * 6 );
    pendB.Reserve( n * sizeof( RgExtRay ) );
    pendC.Reserve( n * sizeof( RgExtHit ) );
    masks.Reserve( n );
    masks2.Reserve( n );
    masksB.Reserve( n );
    capPx = n;
}

size_t PathTracer::OwnBytes()
{
    return 2 * size_t( GRID_ENTRIES ) * 16 + 3 * sizeof( uint64_t );
}

void PathTracer::Init( hipStream_t s )
{
    for( auto& e : traceEv )
    {
        RG_HIP_CHECK( hipEventCreate( &e ) );
    }
    frameAccum.Reserve( 3 * sizeof( uint64_t ) );
    blueNoise = SharedBlueNoise( s );
    for( int k = 0; k < 2; k++ )
    {
        grid[ k ].Reserve( size_t( GRID_ENTRIES ) * 16 );
        // empty cells: selected = LIGHT_INDEX_NONE
        std::vector< uint32_t > init( size_t( GRID_ENTRIES ) * 4, 0u );
        for( size_t e = 0; e < GRID_ENTRIES; e++ )
        {
            init[ e * 4 ] = LIGHT_INDEX_NONE;
        }
        RG_HIP_CHECK( hipMemcpyAsync( grid[ k ].ptr, init.data(), init.size() * 4, hipMemcpyHostToDevice, s ) );
        RG_HIP_CHECK( hipStreamSynchronize( s ) );
    }
}

void PathTracer::Resize( uint32_t w, uint32_t h )
{
    if( w == W && h == H )
    {
        return;
    }
    W              = w;
    H              = h;
    const size_t n = size_t( w ) * h;
    if( wf == &own )
    {
        own.Reserve( n );
    }
    RG_REQUIRE( wf->capPx >= n, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render size exceeds the shared wavefront buffers" );
}

void PathTracer::UploadGeomMap( const std::vector< uint32_t >& map, hipStream_t s )
{
    geomMapHost = map; // stays alive until the frame's stream is synchronised
    geomMap.Reserve( geomMapHost.size() * 4 + 4 );
    geomMapCount = uint32_t( geomMapHost.size() );
    RG_HIP_CHECK( hipMemcpyAsync( geomMap.ptr, geomMapHost.data(), geomMapHost.size() * 4, hipMemcpyHostToDevice, s ) );
}

void PathTracer::UploadLights( const SceneCollector& sc, hipStream_t s )
{
    lightArraySize = uint32_t( sc.lights.size() );
    lightCount     = lightArraySize - 1;
    lights.Reserve( sc.lights.size() * sizeof( LightEncoded ) );
    prevToCur.Reserve( sc.prevToCur.size() * 4 + 4 );
    prevToCurCount = uint32_t( sc.prevToCur.size() );
    RG_HIP_CHECK( hipMemcpyAsync( lights.ptr, sc.lights.data(), sc.lights.size() * sizeof( LightEncoded ), hipMemcpyHostToDevice, s ) );
    RG_HIP_CHECK( hipMemcpyAsync( prevToCur.ptr, sc.prevToCur.data(), sc.prevToCur.size() * 4, hipMemcpyHostToDevice, s ) );
}

uint64_t PathTracer::Trace( const RgExtFrameParams& p, const BvhObject& bvh, TraceContext& tc, const DeviceScene& scene,
                            FramebufferSet& fb, uint32_t traceFlags, hipStream_t s, FrameTimings* timing )
{
    ( void )timing;
    TraceSlot one{ this, &p, &fb };
    return TraceFrames( &one, 1, bvh, tc, &scene, *wf, traceFlags, s );
}

bool GroupableScene( const DeviceScene& scene )
{
    for( uint32_t c = 0; c < GC_COUNT; c++ )
    {
        if( c != GC_WORLD_0 && scene.classTris[ c ] != 0 )
        {
            return false;
        }
    }
    return !scene.alphaTested && !scene.reflRefrGeometry;
}

bool TraceCompatible( const RgExtFrameParams& a, const RgExtFrameParams& b )
{
    return PrimaryCullMask( a ) == PrimaryCullMask( b ) && a.rayCullMaskWorld == b.rayCullMaskWorld &&
           a.reflectRefractMaxDepth == b.reflectRefractMaxDepth && a.rayCullBackFaces == b.rayCullBackFaces &&
           a.directionalLightExists == b.directionalLightExists && ( a.lightCount > 0 ) == ( b.lightCount > 0 );
}

uint64_t TraceFrames( const TraceSlot* slots, uint32_t B, const BvhObject& bvh, TraceContext& tc, const DeviceScene* sharedScene, Wavefront& wf,
                      uint32_t traceFlags, hipStream_t s )
{
    RG_REQUIRE( B >= 1 && B <= MAX_TRACE_SLOTS, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "frames per batch out of range" );
    const RgExtFrameParams& p0      = *slots[ 0 ].p;
    const bool              grouped = sharedScene == nullptr;
    for( uint32_t k = 1; k < B; k++ )
    {
        RG_REQUIRE( TraceCompatible( p0, *slots[ k ].p ), RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                    "frames of one batch differ in their ray masks, reflection depth or lights" );
    }
    if( grouped )
    {
        RG_REQUIRE( bvh.sceneCount > 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "frames of several scenes need their grouped BVH" );
        for( uint32_t k = 0; k < B; k++ )
        {
            RG_REQUIRE( slots[ k ].scene && GroupableScene( *slots[ k ].scene ) && slots[ k ].group < bvh.sceneCount &&
                            ( k == 0 || slots[ k ].group >= slots[ k - 1 ].group ),
                        RG_RESULT_WRONG_FUNCTION_ARGUMENT, "grouped frames need plain scenes of the group in group order" );
            // every ray mask then admits the one class there is
            RG_REQUIRE( ( slots[ k ].p->rayCullMaskWorld & slots[ k ].p->rayCullMaskWorldShadow & IMASK_WORLD_0 ) != 0u,
                        RG_RESULT_WRONG_FUNCTION_ARGUMENT, "grouped frames must see world 0 with every ray" );
        }
    }
    // Each frame's rays, hits and pending records are one slice of the
    // wavefront: frame k's pixel i is entry off[k] + i, so a traversal
    // launch covers every frame's rays of a stage at once.
    size_t   off[ MAX_TRACE_SLOTS ], offG[ MAX_TRACE_SLOTS ];
    uint32_t nPx[ MAX_TRACE_SLOTS ], gw[ MAX_TRACE_SLOTS ], gh[ MAX_TRACE_SLOTS ];
    size_t   total = 0, totalG = 0;
    for( uint32_t k = 0; k < B; k++ )
    {
        PathTracer& pt = *slots[ k ].pt;
        pt.Resize( slots[ k ].p->renderWidth, slots[ k ].p->renderHeight );
        nPx[ k ] = pt.W * pt.H;
        gw[ k ]  = ( pt.W + 1 ) / 3 > 1 ? ( pt.W + 1 ) / 3 : 1;
        gh[ k ]  = ( pt.H + 1 ) / 3 > 1 ? ( pt.H + 1 ) / 3 : 1;
        off[ k ] = total;
        offG[ k ] = totalG;
        total += nPx[ k ];
        totalG += size_t( gw[ k ] ) * gh[ k ];
    }
    RG_REQUIRE( total <= wf.capPx, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "the batch's pixels exceed the wavefront capacity" );
    PathTracer&        lead  = *slots[ 0 ].pt; // times and counts the batch's traversal launches
    const DeviceScene& scene = grouped ? *slots[ 0 ].scene : *sharedScene; // scene-wide switches (equal in a group)
    const TexView      tex   = scene.tex;
    const AlphaRec*    alpha = !grouped && scene.alphaTested ? scene.alphaRecs.as< AlphaRec >() : nullptr;
    RgExtRay* R  = wf.rays.as< RgExtRay >();
    RgExtHit* Hh = wf.hits.as< RgExtHit >();
    RgExtRay* R2 = wf.rays2.as< RgExtRay >();
    RgExtHit* H2 = wf.hits2.as< RgExtHit >();
    float4*   PA = wf.pendA.as< float4 >();
    RgExtRay* PB = wf.pendB.as< RgExtRay >();
    RgExtHit* PC = wf.pendC.as< RgExtHit >();
    uint8_t*  M  = wf.masks.as< uint8_t >();
    uint8_t*  M2 = wf.masks2.as< uint8_t >();
    uint8_t*  MB = wf.masksB.as< uint8_t >();
    uint64_t  traced = 0;
    const bool haveGeometry = grouped ? bvh.nodeCount > 0 : scene.triCount > 0 && !bvh.frameSegs.empty();

    lead.traceLaunches = 0;
    RG_HIP_CHECK( hipMemsetAsync( lead.frameAccum.ptr, 0, 3 * sizeof( uint64_t ), s ) );
    TraceAccum accum;
    accum.stats  = lead.frameAccum.as< uint64_t >();
    accum.errors = reinterpret_cast< uint32_t* >( lead.frameAccum.as< uint64_t >() + 2 );
    // The frame's TLAS: each geometry class with its instance mask, or 0
    // when rayCullMaskWorld leaves its world out (ASManager.cpp:837-963).
    uint32_t classMasks[ GC_COUNT ];
    for( uint32_t c = 0; c < GC_COUNT; c++ )
    {
        classMasks[ c ] = GeomClassPresent( c, p0.rayCullMaskWorld ) ? GeomClassMask( c ) : 0u;
    }
    // rr's rays carry their cull masks in mm (null: every ray has mask);
    // frame k's rays start at offs[k].  A grouped launch gives each scene of
    // the group the rays of its frames (plain scenes: the masks admit every
    // triangle, null rays have an empty interval).
    std::vector< uint32_t > groupOffs;
    auto trace = [ & ]( RgExtRay* rr, const uint8_t* mm, uint32_t mask, RgExtHit* hh, uint32_t flags, size_t count, const size_t* offs ) {
        if( haveGeometry )
        {
            const uint32_t k  = lead.traceLaunches < PathTracer::TRACE_EVENT_PAIRS ? lead.traceLaunches : PathTracer::TRACE_EVENT_PAIRS - 1;
            const uint32_t fl = flags | traceFlags | ( lead.collectStats ? RG_EXT_TRACE_COLLECT_STATS : 0u );
            if( grouped )
            {
                groupOffs.assign( bvh.sceneCount + 1, uint32_t( count ) );
                uint32_t f = 0;
                for( uint32_t g = 0; g < bvh.sceneCount; g++ )
                {
                    while( f < B && slots[ f ].group < g )
                    {
                        f++;
                    }
                    groupOffs[ g ] = f < B ? uint32_t( offs[ f ] ) : uint32_t( count );
                }
                TraceRaysBatch( bvh, tc, rr, groupOffs.data(), hh, fl, s, lead.traceEv[ 2 * k ], lead.traceEv[ 2 * k + 1 ], &accum );
            }
            else
            {
                TraceRaysFrame( bvh, tc, rr, mm, mask, uint32_t( count ), hh, fl, classMasks, GC_COUNT, s, lead.traceEv[ 2 * k ],
                                lead.traceEv[ 2 * k + 1 ], &accum, alpha, &tex );
            }
            lead.traceLaunches = k + 1;
        }
        else
        {
            // no geometry: every ray misses
            RG_HIP_CHECK( hipMemsetAsync( hh, 0xFF, count * sizeof( RgExtHit ), s ) );
        }
        traced += count;
    };
    struct Frame
    {
        const RgExtFrameParams* p;
        PathTracer*             pt;
        const DeviceScene*      sc;
        SceneView               sv;
        FbPtrs                  P;
        LightView               lv;
        dim3                    grid2;
        size_t                  o, n;
    };
    Frame f[ MAX_TRACE_SLOTS ];
    for( uint32_t k = 0; k < B; k++ )
    {
        PathTracer& pt = *slots[ k ].pt;
        f[ k ].p     = slots[ k ].p;
        f[ k ].pt    = &pt;
        f[ k ].sc    = grouped ? slots[ k ].scene : sharedScene;
        const DeviceScene& sk = *f[ k ].sc;
        f[ k ].sv    = SceneView{ sk.tris.as< float >(), sk.triNormals.as< float >(), sk.triGeom.as< uint32_t >(), sk.geoms.as< GeomRecord >(),
                                  sk.vertices.as< RgPrimitiveVertex >(), sk.indices.as< uint32_t >(), sk.tex };
        f[ k ].P     = slots[ k ].fb->Ptrs();
        f[ k ].lv    = LightView{ pt.lights.as< LightEncoded >(), pt.lightArraySize, pt.prevToCur.as< uint32_t >(), pt.prevToCurCount,
                                  pt.grid[ pt.gridParity ].as< uint4 >(), pt.grid[ pt.gridParity ^ 1 ].as< uint4 >(), pt.blueNoise };
        f[ k ].grid2 = dim3( DivUp( pt.W, BX ), DivUp( pt.H, BY ) );
        f[ k ].o     = off[ k ];
        f[ k ].n     = nPx[ k ];
    }
    auto each = [ & ]( auto&& fn ) {
        for( uint32_t k = 0; k < B; k++ )
        {
            fn( f[ k ] );
        }
    };

    // primary
    each( [ & ]( Frame& q ) {
        k_primary_rays<<< q.grid2, THREADS, 0, s >>>( *q.p, R + q.o );
        CheckLaunch( "k_primary_rays" );
    } );
    trace( R, nullptr, PrimaryCullMask( p0 ), Hh, RG_EXT_TRACE_CLOSEST, total, off );
    each( [ & ]( Frame& q ) {
        ( q.sv.tex.count > 0 ? k_primary_shade< true > : k_primary_shade< false > )<<< q.grid2, THREADS, 0, s >>>( *q.p, Hh + q.o, q.sv, q.P );
        CheckLaunch( "k_primary_shade" );
    } );

    // reflections and refractions replace the G-buffer surface; without
    // such geometry the reference's loop ends at depth 0 with no writes
    if( p0.reflectRefractMaxDepth > 0 && haveGeometry && scene.reflRefrGeometry )
    {
        for( uint32_t d = 0; d < p0.reflectRefractMaxDepth && d < 16u; d++ )
        {
            each( [ & ]( Frame& q ) {
                RrState st{ PA + 6 * q.o, q.n };
                ( q.sv.tex.count > 0 ? k_reflrefr_a< true > : k_reflrefr_a< false > )<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, q.sv, R + q.o, M + q.o,
                                                                                                                Hh + q.o, st, d );
                CheckLaunch( "k_reflrefr_a" );
            } );
            trace( R, M, 0u, Hh, RG_EXT_TRACE_CLOSEST, total, off );
            each( [ & ]( Frame& q ) {
                RrState st{ PA + 6 * q.o, q.n };
                ( q.sv.tex.count > 0 ? k_reflrefr_b< true > : k_reflrefr_b< false > )<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, q.sv, R + q.o, Hh + q.o,
                                                                                                                st, d );
                CheckLaunch( "k_reflrefr_b" );
            } );
        }
    }

    // light grid (current grid = gridParity, previous = the other one)
    each( [ & ]( Frame& q ) {
        PathTracer& pt = *q.pt;
        if( q.p->lightCount > 0 )
        {
            q.lv.grid     = pt.grid[ pt.gridParity ^ 1 ].as< uint4 >();
            q.lv.gridPrev = pt.grid[ pt.gridParity ].as< uint4 >();
            k_light_grid<<< DivUp( GRID_ENTRIES, THREADS ), THREADS, 0, s >>>( *q.p, q.lv, pt.grid[ pt.gridParity ^ 1 ].as< uint4 >() );
            CheckLaunch( "k_light_grid" );
            pt.gridParity ^= 1;
            q.lv.grid     = pt.grid[ pt.gridParity ].as< uint4 >();
            q.lv.gridPrev = pt.grid[ pt.gridParity ^ 1 ].as< uint4 >();
        }
    } );

    // initial reservoirs (directional visibility)
    each( [ & ]( Frame& q ) {
        k_initial_a<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, q.lv, R + q.o, M + q.o, reinterpret_cast< uint4* >( PA + 6 * q.o ) );
        CheckLaunch( "k_initial_a" );
    } );
    if( p0.directionalLightExists )
    {
        // without a directional light every ray is null and k_initial_b reads no hit
        trace( R, M, 0u, Hh, RG_EXT_TRACE_ANY_HIT, total, off );
    }
    each( [ & ]( Frame& q ) {
        k_initial_b<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, Hh + q.o, reinterpret_cast< uint4* >( PA + 6 * q.o ) );
        CheckLaunch( "k_initial_b" );
    } );

    // direct illumination
    each( [ & ]( Frame& q ) {
        k_direct_a<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, q.lv, R + q.o, M + q.o, PA + 6 * q.o, PA + 6 * q.o + q.n );
        CheckLaunch( "k_direct_a" );
    } );
    trace( R, M, 0u, Hh, RG_EXT_TRACE_ANY_HIT, total, off );
    each( [ & ]( Frame& q ) {
        k_direct_b<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, R + q.o, Hh + q.o, PA + 6 * q.o, PA + 6 * q.o + q.n );
        CheckLaunch( "k_direct_b" );
    } );

    // indirect: bounce 1, then the shadow ray at its hit and bounce 2
    each( [ & ]( Frame& q ) {
        k_indirect_a<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P, q.lv, R + q.o, M + q.o, IndirPend{ PA + 6 * q.o, q.n } );
        CheckLaunch( "k_indirect_a" );
    } );
    trace( R, M, 0u, Hh, RG_EXT_TRACE_CLOSEST, total, off );
    each( [ & ]( Frame& q ) {
        ( q.sv.tex.count > 0 ? k_indirect_b< true > : k_indirect_b< false > )<<< q.grid2, THREADS, 0, s >>>(
            *q.p, q.P, q.sv, q.lv, R + q.o, Hh + q.o, PB + q.o, MB + q.o, R2 + q.o, M2 + q.o, IndirPend{ PA + 6 * q.o, q.n } );
        CheckLaunch( "k_indirect_b" );
    } );
    trace( PB, MB, 0u, PC, RG_EXT_TRACE_ANY_HIT, total, off );
    trace( R2, M2, 0u, H2, RG_EXT_TRACE_CLOSEST, total, off );
    each( [ & ]( Frame& q ) {
        ( q.sv.tex.count > 0 ? k_indirect_c< true > : k_indirect_c< false > )<<< q.grid2, THREADS, 0, s >>>(
            *q.p, q.P, q.sv, q.lv, PC + q.o, R2 + q.o, H2 + q.o, Hh + q.o, IndirPend{ PA + 6 * q.o, q.n } );
        CheckLaunch( "k_indirect_c" );
        k_indirect_final<<< q.grid2, THREADS, 0, s >>>( *q.p, q.P );
        CheckLaunch( "k_indirect_final" );
    } );

    // gradient samples (CalculateGradientsSamples, PathTracer.cpp:196-219):
    // frame k's strata are entries offG[k] .. of the ray and hit arrays
    for( uint32_t k = 0; k < B; k++ )
    {
        Frame&     q = f[ k ];
        const dim3 gridG( DivUp( gw[ k ], BX ), DivUp( gh[ k ], BY ) );
        float4*    lum = PA + 6 * q.o;
        GradPend   gp{ lum, reinterpret_cast< uint4* >( lum + size_t( gw[ k ] ) * gh[ k ] ) };
        k_grad_a<<< gridG, THREADS, 0, s >>>( *q.p, q.P, q.sv, q.lv, q.pt->geomMap.as< uint32_t >(), q.pt->geomMapCount, q.sc->geomCount,
                                             R + offG[ k ], M + offG[ k ], gp );
        CheckLaunch( "k_grad_a" );
    }
    trace( R, M, 0u, Hh, RG_EXT_TRACE_ANY_HIT, totalG, offG );
    for( uint32_t k = 0; k < B; k++ )
    {
        Frame&     q = f[ k ];
        const dim3 gridG( DivUp( gw[ k ], BX ), DivUp( gh[ k ], BY ) );
        float4*    lum = PA + 6 * q.o;
        GradPend   gp{ lum, reinterpret_cast< uint4* >( lum + size_t( gw[ k ] ) * gh[ k ] ) };
        k_grad_b<<< gridG, THREADS, 0, s >>>( *q.p, q.P, Hh + offG[ k ], gp );
        CheckLaunch( "k_grad_b" );
    }
    return traced;
}

} // namespace rg
