// Scene collection (host) and world-space triangle assembly (device).
//
// Follows Scene::UploadPrimitive / VertexCollector::AddPrimitive
// (Source/Scene.cpp:106-145, Source/VertexCollector.cpp:130-342): the
// caller's arrays are copied before the call returns; transforms are applied
// on the device (the reference lets the AS build apply transformData,
// VertexCollector.cpp:259-261, and getTriangle applies them for shading,
// Shaders/VertexData.inl:317-493).
#include "rg_device_math.h"
#include "rg_scene.h"
#include "rg_textures.h"

#include <string>

#include <algorithm>
#include <thread>
#include <vector>

#include <cmath>

namespace rg
{

namespace
{
// Row-major 3x4 affine inverse in FP64.
void InvertAffine( const float m[ 12 ], double out[ 12 ] )
{
    const double a = m[ 0 ], b = m[ 1 ], c = m[ 2 ], d = m[ 4 ], e = m[ 5 ], f = m[ 6 ], g = m[ 8 ], h = m[ 9 ], i = m[ 10 ];
    const double det = a * ( e * i - f * h ) - b * ( d * i - f * g ) + c * ( d * h - e * g );
    const double inv = det != 0.0 ? 1.0 / det : 0.0;
    double       r[ 9 ] = { ( e * i - f * h ) * inv, ( c * h - b * i ) * inv, ( b * f - c * e ) * inv,
                            ( f * g - d * i ) * inv, ( a * i - c * g ) * inv, ( c * d - a * f ) * inv,
                            ( d * h - e * g ) * inv, ( b * g - a * h ) * inv, ( a * e - b * d ) * inv };
    for( int row = 0; row < 3; row++ )
    {
        out[ row * 4 + 0 ] = r[ row * 3 + 0 ];
        out[ row * 4 + 1 ] = r[ row * 3 + 1 ];
        out[ row * 4 + 2 ] = r[ row * 3 + 2 ];
        out[ row * 4 + 3 ] = -( r[ row * 3 + 0 ] * m[ 3 ] + r[ row * 3 + 1 ] * m[ 7 ] + r[ row * 3 + 2 ] * m[ 11 ] );
    }
}

// out = A (3x4) * B (3x4) as affine maps.
void ComposeAffine( const float A[ 12 ], const double B[ 12 ], float out[ 12 ] )
{
    for( int r = 0; r < 3; r++ )
    {
        for( int c = 0; c < 4; c++ )
        {
            double v = double( A[ r * 4 + 0 ] ) * B[ 0 * 4 + c ] + double( A[ r * 4 + 1 ] ) * B[ 1 * 4 + c ] +
                       double( A[ r * 4 + 2 ] ) * B[ 2 * 4 + c ];
            if( c == 3 )
            {
                v += A[ r * 4 + 3 ];
            }
            out[ r * 4 + c ] = float( v );
        }
    }
}

float UnpackChannel( uint32_t c, int i )
{
    return float( ( c >> ( 8 * i ) ) & 0xFFu ) / 255.0f;
}

float Saturate( float x )
{
    return x < 0.0f ? 0.0f : ( x > 1.0f ? 1.0f : x );
}
} // namespace

void SceneCollector::BeginFrame()
{
    vertices.Clear();
    indices.Clear();
    geoms.Clear();
    for( uint32_t c = 0; c < GC_COUNT; c++ )
    {
        classGeoms[ c ].clear();
        classUids[ c ].clear();
        classTris[ c ] = 0;
    }
    triCount = 0;
    sealed   = false;
    prevModels.swap( curModels );
    curModels.clear();
    prevGeomUids.swap( geomUids );
    geomUids.clear();

    prevLightIndex.clear();
    for( uint32_t i = 0; i < lights.size(); i++ )
    {
        if( lights[ i ].lightType != LIGHT_TYPE_NONE )
        {
            prevLightIndex[ lightIds[ i ] ] = i;
        }
    }
    prevLightCount = uint32_t( lights.size() );
    lights.assign( 1, LightEncoded{} ); // slot 0: directional (LIGHT_ARRAY_DIRECTIONAL_LIGHT_OFFSET)
    lightIds.assign( 1, 0 );
    hasDirectional = false;
}

namespace
{
// Copies into pinned staging: a large primitive (a 524k-triangle terrain is
// 100 MB of vertices) is split over a few threads, since one thread's
// memcpy runs at a fraction of the host's memory bandwidth.
void StagingCopy( void* dst, const void* src, size_t bytes )
{
    constexpr size_t SPLIT = size_t( 8 ) << 20;
    const unsigned   hw    = std::thread::hardware_concurrency();
    const unsigned   T     = bytes >= SPLIT ? std::min( 8u, hw > 1 ? hw : 1u ) : 1u;
    if( T <= 1 )
    {
        std::memcpy( dst, src, bytes );
        return;
    }
    const size_t             chunk = ( bytes / T + 63 ) & ~size_t( 63 );
    std::vector< std::thread > pool;
    for( unsigned t = 1; t < T; t++ )
    {
        const size_t off = chunk * t;
        if( off >= bytes )
        {
            break;
        }
        const size_t len = std::min( chunk, bytes - off );
        pool.emplace_back( [ = ] { std::memcpy( static_cast< char* >( dst ) + off, static_cast< const char* >( src ) + off, len ); } );
    }
    std::memcpy( dst, src, std::min( chunk, bytes ) );
    for( auto& th : pool )
    {
        th.join();
    }
}
} // namespace

void SceneCollector::AddPrimitive( const RgMeshInfo& mesh, const RgMeshPrimitiveInfo& prim )
{
    if( prim.vertexCount == 0 || prim.pVertices == nullptr )
    {
        return; // VulkanDevice.cpp:935-938
    }
    // Rasterised primitives (sky, translucent) are not ray traced
    // (IsRasterized, VulkanDevice.cpp:899-923).
    const bool media = ( prim.flags & ( RG_MESH_PRIMITIVE_GLASS | RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_ACID ) ) != 0;
    if( ( prim.flags & RG_MESH_PRIMITIVE_SKY ) ||
        ( !media && ( ( prim.flags & RG_MESH_PRIMITIVE_TRANSLUCENT ) || UnpackChannel( prim.color, 3 ) < 0.98f ) ) )
    {
        return;
    }
    uint32_t tris = 0;
    if( prim.pIndices && prim.indexCount > 0 )
    {
        RG_REQUIRE( prim.indexCount % 3 == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "indexCount must be a multiple of 3" );
        tris = prim.indexCount / 3;
        for( uint32_t k = 0; k < prim.indexCount; k++ )
        {
            RG_REQUIRE( prim.pIndices[ k ] < prim.vertexCount, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "index out of the vertex range" );
        }
    }
    else
    {
        tris = prim.vertexCount / 3;
    }
    if( tris == 0 )
    {
        return;
    }
    RG_REQUIRE( uint64_t( triCount ) + tris < ( 1ull << 28 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "too many triangles in one frame" );

    RG_REQUIRE( !sealed, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "primitive uploaded after the frame's geometry was sealed" );
    const uint32_t cls = GeomClassOf( prim.flags );

    GeomRecord g{};
    std::memcpy( g.model, &mesh.transform.matrix[ 0 ][ 0 ], sizeof( g.model ) );
    g.baseVertex = uint32_t( vertices.size );
    g.baseIndex  = 0xFFFFFFFFu;
    g.firstTri   = classTris[ cls ]; // within its class until Seal()
    g.triCount   = tris;
    g.color      = prim.color;
    g.emissive   = Saturate( prim.emissive );
    const bool pbr = prim.pEditorInfo && prim.pEditorInfo->pbrInfoExists;
    g.roughness    = pbr ? Saturate( prim.pEditorInfo->pbrInfo.roughnessDefault ) : 1.0f;
    g.metallic     = pbr ? Saturate( prim.pEditorInfo->pbrInfo.metallicDefault ) : 0.0f;
    g.flags        = prim.flags;
    if( textures && prim.pTextureName && prim.pTextureName[ 0 ] )
    {
        const std::string base( prim.pTextureName );
        g.texture     = textures->Find( base.c_str() );
        g.texOrm      = textures->Find( ( base + "_orm" ).c_str() );
        g.texNormal   = textures->Find( ( base + "_n" ).c_str() );
        g.texEmissive = textures->Find( ( base + "_e" ).c_str() );
    }

    // Previous transform of the same (object, primitive): UniqueID::MakeForPrimitive.
    const uint64_t uid = ( uint64_t( mesh.uniqueObjectID ) << 32 ) | prim.primitiveIndexInMesh;
    std::array< float, 12 > cur;
    std::memcpy( cur.data(), g.model, sizeof( g.model ) );
    curModels[ uid ] = cur;
    auto prev        = prevModels.find( uid );
    if( prev != prevModels.end() && std::memcmp( prev->second.data(), g.model, sizeof( g.model ) ) != 0 )
    {
        double inv[ 12 ];
        InvertAffine( g.model, inv );
        ComposeAffine( prev->second.data(), inv, g.prevFromCur );
        g.isStatic = 0;
    }
    else
    {
        for( int k = 0; k < 12; k++ )
        {
            g.prevFromCur[ k ] = ( k % 5 == 0 ) ? 1.0f : 0.0f;
        }
        g.isStatic = 1;
    }

    StagingCopy( vertices.Append( prim.vertexCount ), prim.pVertices, size_t( prim.vertexCount ) * sizeof( RgPrimitiveVertex ) );
    if( prim.pIndices && prim.indexCount > 0 )
    {
        g.baseIndex = uint32_t( indices.size );
        StagingCopy( indices.Append( prim.indexCount ), prim.pIndices, size_t( prim.indexCount ) * 4 );
    }
    if( cls == GC_WORLD_0 )
    {
        *geoms.Append( 1 ) = g;
        geomUids.push_back( uid );
    }
    else
    {
        classGeoms[ cls ].push_back( g );
        classUids[ cls ].push_back( uid );
    }
    classTris[ cls ] += tris;
    triCount += tris;
}

void SceneCollector::Seal()
{
    if( sealed )
    {
        return;
    }
    sealed        = true;
    uint32_t base = classTris[ GC_WORLD_0 ];
    for( uint32_t c = GC_WORLD_0 + 1; c < GC_COUNT; c++ )
    {
        for( GeomRecord g : classGeoms[ c ] )
        {
            g.firstTri += base;
            *geoms.Append( 1 ) = g;
        }
        geomUids.insert( geomUids.end(), classUids[ c ].begin(), classUids[ c ].end() );
        classGeoms[ c ].clear();
        classUids[ c ].clear();
        base += classTris[ c ];
    }
}

// Light encoders: LightManager.cpp:99-234.
float SceneCollector::LightStyle( const RgLightExtraInfo& extra ) const
{
    // an index outside the table counts as no style
    if( extra.exists && extra.lightstyle >= 0 && size_t( extra.lightstyle ) < lightstyles.size() )
    {
        return lightstyles[ size_t( extra.lightstyle ) ];
    }
    return 1.0f;
}

void SceneCollector::AddDirectional( const RgDirectionalLightUploadInfo& info )
{
    if( hasDirectional )
    {
        return; // "Only one directional light is allowed" (LightManager.cpp:432-436)
    }
    if( info.intensity <= 0.00001f || ( info.color & 0x00FFFFFFu ) == 0 || info.angularDiameterDegrees < 0.0f )
    {
        return;
    }
    float       d[ 3 ] = { info.direction.data[ 0 ], info.direction.data[ 1 ], info.direction.data[ 2 ] };
    const float len    = std::sqrt( d[ 0 ] * d[ 0 ] + d[ 1 ] * d[ 1 ] + d[ 2 ] * d[ 2 ] );
    const float  mult = LightStyle( info.extra );
    LightEncoded lt{};
    lt.lightType = LIGHT_TYPE_DIRECTIONAL;
    for( int i = 0; i < 3; i++ )
    {
        lt.color[ i ] = UnpackChannel( info.color, i ) * info.intensity * mult;
        lt.data0[ i ] = len > 0.0f ? d[ i ] / len : 0.0f;
    }
    lt.data0[ 3 ]  = 0.5f * info.angularDiameterDegrees * float( M_PI ) / 180.0f;
    lights[ 0 ]    = lt;
    lightIds[ 0 ]  = info.uniqueID;
    hasDirectional = true;
}

void SceneCollector::AddSphere( const RgSphericalLightUploadInfo& info )
{
    if( info.intensity <= 0.00001f || ( info.color & 0x00FFFFFFu ) == 0 )
    {
        return;
    }
    const float  radius = info.radius > 0.005f ? info.radius : 0.005f; // MIN_SPHERE_RADIUS
    const float  area   = float( M_PI ) * radius * radius;
    const float  mult = LightStyle( info.extra );
    LightEncoded lt{};
    lt.lightType = LIGHT_TYPE_SPHERE;
    for( int i = 0; i < 3; i++ )
    {
        lt.color[ i ] = UnpackChannel( info.color, i ) * info.intensity / area * mult;
        lt.data0[ i ] = info.position.data[ i ];
    }
    lt.data0[ 3 ] = radius;
    lights.push_back( lt );
    lightIds.push_back( info.uniqueID );
}

void SceneCollector::AddSpot( const RgSpotLightUploadInfo& info )
{
    if( info.intensity <= 0.00001f || ( info.color & 0x00FFFFFFu ) == 0 || info.radius < 0.0f || info.angleOuter <= 0.0f )
    {
        return;
    }
    const float radius = info.radius > 0.005f ? info.radius : 0.005f;
    const float area   = float( M_PI ) * radius * radius;
    float       d[ 3 ] = { info.direction.data[ 0 ], info.direction.data[ 1 ], info.direction.data[ 2 ] };
    const float len    = std::sqrt( d[ 0 ] * d[ 0 ] + d[ 1 ] * d[ 1 ] + d[ 2 ] * d[ 2 ] );
    const float deg    = float( M_PI ) / 180.0f;
    auto        clampc = [ & ]( float a ) { return a < 0.0f ? 0.0f : ( a > 89.0f * deg ? 89.0f * deg : a ); };
    const float inner  = std::min( info.angleInner, info.angleOuter - 1.0f * deg );
    const float  mult = LightStyle( info.extra );
    LightEncoded lt{};
    lt.lightType = LIGHT_TYPE_SPOT;
    for( int i = 0; i < 3; i++ )
    {
        lt.color[ i ] = UnpackChannel( info.color, i ) * info.intensity / area * mult;
        lt.data0[ i ] = info.position.data[ i ];
        lt.data1[ i ] = len > 0.0f ? d[ i ] / len : 0.0f;
    }
    lt.data0[ 3 ] = radius;
    lt.data2[ 0 ] = std::cos( clampc( inner ) );
    lt.data2[ 1 ] = std::cos( clampc( info.angleOuter ) );
    lights.push_back( lt );
    lightIds.push_back( info.uniqueID );
}

void SceneCollector::AddPolygonal( const RgPolygonalLightUploadInfo& info )
{
    if( info.intensity <= 0.00001f || ( info.color & 0x00FFFFFFu ) == 0 )
    {
        return;
    }
    const float* p0 = info.positions[ 0 ].data;
    const float* p1 = info.positions[ 1 ].data;
    const float* p2 = info.positions[ 2 ].data;
    const float  e1[ 3 ] = { p1[ 0 ] - p0[ 0 ], p1[ 1 ] - p0[ 1 ], p1[ 2 ] - p0[ 2 ] };
    const float  e2[ 3 ] = { p2[ 0 ] - p0[ 0 ], p2[ 1 ] - p0[ 1 ], p2[ 2 ] - p0[ 2 ] };
    const float  n[ 3 ]  = { e1[ 1 ] * e2[ 2 ] - e1[ 2 ] * e2[ 1 ], e1[ 2 ] * e2[ 0 ] - e1[ 0 ] * e2[ 2 ], e1[ 0 ] * e2[ 1 ] - e1[ 1 ] * e2[ 0 ] };
    const float  len     = std::sqrt( n[ 0 ] * n[ 0 ] + n[ 1 ] * n[ 1 ] + n[ 2 ] * n[ 2 ] );
    if( !( len > 0.0f ) )
    {
        return;
    }
    const float  area = len * 0.5f;
    const float  mult = LightStyle( info.extra );
    LightEncoded lt{};
    lt.lightType = LIGHT_TYPE_TRIANGLE;
    for( int i = 0; i < 3; i++ )
    {
        lt.color[ i ] = UnpackChannel( info.color, i ) * info.intensity / area * mult;
        lt.data0[ i ] = p0[ i ];
        lt.data1[ i ] = p1[ i ];
        lt.data2[ i ] = p2[ i ];
    }
    lt.data0[ 3 ] = n[ 0 ];
    lt.data1[ 3 ] = n[ 1 ];
    lt.data2[ 3 ] = n[ 2 ];
    lights.push_back( lt );
    lightIds.push_back( info.uniqueID );
}

void SceneCollector::FinishLights()
{
    // lightSources_Index_PrevToCur: previous index -> current index, or
    // UINT32_MAX when the light is gone (LightManager.cpp FillMatchPrev).
    prevToCur.assign( prevLightCount > 0 ? prevLightCount : 1, 0xFFFFFFFFu );
    for( uint32_t i = 0; i < lights.size(); i++ )
    {
        if( lights[ i ].lightType == LIGHT_TYPE_NONE )
        {
            continue;
        }
        auto it = prevLightIndex.find( lightIds[ i ] );
        if( it != prevLightIndex.end() && it->second < prevToCur.size() )
        {
            prevToCur[ it->second ] = i;
        }
    }
}

uint32_t SceneCollector::LightIndexOf( const uint64_t* uniqueId ) const
{
    if( !uniqueId )
    {
        return LIGHT_INDEX_NONE;
    }
    for( uint32_t i = 0; i < lights.size(); i++ )
    {
        if( lights[ i ].lightType != LIGHT_TYPE_NONE && lightIds[ i ] == *uniqueId )
        {
            return i;
        }
    }
    return LIGHT_INDEX_NONE;
}

void SceneCollector::GeomPrevToCur( std::vector< uint32_t >& out )
{
    Seal();
    std::unordered_map< uint64_t, uint32_t > cur;
    cur.reserve( geomUids.size() );
    for( uint32_t i = 0; i < geomUids.size(); i++ )
    {
        cur.emplace( geomUids[ i ], i ); // keeps the first of duplicate ids
    }
    out.assign( prevGeomUids.size() > 0 ? prevGeomUids.size() : 1, 0xFFFFFFFFu );
    for( uint32_t i = 0; i < prevGeomUids.size(); i++ )
    {
        auto it = cur.find( prevGeomUids[ i ] );
        if( it != cur.end() )
        {
            out[ i ] = it->second;
        }
    }
}

// ------------------------------------------------------------ device

namespace
{
constexpr int THREADS = 256;

__global__ __launch_bounds__( THREADS ) void k_tri_geom( const GeomRecord* geoms, uint32_t geomCount, uint32_t triCount, uint32_t* triGeom )
{
    const uint32_t t = blockIdx.x * THREADS + threadIdx.x;
    if( t >= triCount )
    {
        return;
    }
    uint32_t lo = 0, hi = geomCount; // last geometry with firstTri <= t
    while( hi - lo > 1 )
    {
        const uint32_t mid = ( lo + hi ) / 2;
        if( geoms[ mid ].firstTri <= t )
        {
            lo = mid;
        }
        else
        {
            hi = mid;
        }
    }
    triGeom[ t ] = lo;
}

__device__ __forceinline__ v3 Xform( const float* m, v3 p, float w )
{
    return mk3( m[ 0 ] * p.x + m[ 1 ] * p.y + m[ 2 ] * p.z + m[ 3 ] * w,
                m[ 4 ] * p.x + m[ 5 ] * p.y + m[ 6 ] * p.z + m[ 7 ] * w,
                m[ 8 ] * p.x + m[ 9 ] * p.y + m[ 10 ] * p.z + m[ 11 ] * w );
}

__global__ __launch_bounds__( THREADS ) void k_assemble( const RgPrimitiveVertex* verts,
                                                         const uint32_t*          indices,
                                                         const GeomRecord*        geoms,
                                                         const uint32_t*          triGeom,
                                                         uint32_t                 triCount,
                                                         float*                   tris,
                                                         float*                   normals,
                                                         AlphaRec*                alpha )
{
    const uint32_t t = blockIdx.x * THREADS + threadIdx.x;
    if( t >= triCount )
    {
        return;
    }
    const GeomRecord& g = geoms[ triGeom[ t ] ];
    const uint32_t    i = t - g.firstTri;
#pragma unroll
    for( int k = 0; k < 3; k++ )
    {
        const uint32_t           idx = g.baseIndex == 0xFFFFFFFFu ? 3 * i + k : indices[ g.baseIndex + 3 * i + k ];
        const RgPrimitiveVertex& v   = verts[ g.baseVertex + idx ];
        const v3 p = Xform( g.model, mk3( v.position[ 0 ], v.position[ 1 ], v.position[ 2 ] ), 1.0f );
        const v3 n = Xform( g.model, mk3( v.normal[ 0 ], v.normal[ 1 ], v.normal[ 2 ] ), 0.0f );
        tris[ size_t( t ) * 9 + 3 * k + 0 ]    = p.x;
        tris[ size_t( t ) * 9 + 3 * k + 1 ]    = p.y;
        tris[ size_t( t ) * 9 + 3 * k + 2 ]    = p.z;
        normals[ size_t( t ) * 9 + 3 * k + 0 ] = n.x;
        normals[ size_t( t ) * 9 + 3 * k + 1 ] = n.y;
        normals[ size_t( t ) * 9 + 3 * k + 2 ] = n.z;
        if( alpha )
        {
            alpha[ t ].uv[ 2 * k + 0 ] = v.texCoord[ 0 ];
            alpha[ t ].uv[ 2 * k + 1 ] = v.texCoord[ 1 ];
        }
    }
    if( alpha )
    {
        // only alpha-tested, textured primitives are tested (RtAlphaTest.rahit)
        alpha[ t ].tex   = ( g.flags & RG_MESH_PRIMITIVE_ALPHA_TESTED ) ? g.texture : TEX_NONE;
        alpha[ t ].color = g.color;
    }
}

// Flat normals (CmVertexPreprocess.comp, VertexPreprocessPartial.inl:43-120):
// a geometry without RG_MESH_PRIMITIVE_DONT_GENERATE_NORMALS
// (GeomInfoManager.cpp:173-176) gets normalize(cross(p1 - p0, p2 - p0)) of
// its local positions written to the vertices of each triangle, triangles
// in index order, so a vertex shared by several triangles keeps the last
// one's normal.  The reference loops over a geometry's triangles in one
// thread; here one thread per triangle, with the last triangle of each
// vertex found first (k_flat_last) so every vertex has one writer.
// GEOM_INST_FLAG_INVERTED_NORMALS is never set by the reference's
// GeomInfoManager, so the sign is +1.
__device__ __forceinline__ bool GeneratesNormals( const GeomRecord& g )
{
    return ( g.flags & RG_MESH_PRIMITIVE_DONT_GENERATE_NORMALS ) == 0u;
}

__global__ __launch_bounds__( THREADS ) void k_flat_last( const uint32_t* indices, const GeomRecord* geoms, const uint32_t* triGeom,
                                                          uint32_t triCount, uint32_t* lastTri )
{
    const uint32_t t = blockIdx.x * THREADS + threadIdx.x;
    if( t >= triCount )
    {
        return;
    }
    const GeomRecord& g = geoms[ triGeom[ t ] ];
    if( !GeneratesNormals( g ) || g.baseIndex == 0xFFFFFFFFu )
    {
        return;
    }
    const uint32_t i = t - g.firstTri;
#pragma unroll
    for( int k = 0; k < 3; k++ )
    {
        atomicMax( &lastTri[ g.baseVertex + indices[ g.baseIndex + 3 * i + k ] ], i + 1u );
    }
}

__global__ __launch_bounds__( THREADS ) void k_flat_write( RgPrimitiveVertex* verts, const uint32_t* indices, const GeomRecord* geoms,
                                                           const uint32_t* triGeom, uint32_t triCount, const uint32_t* lastTri )
{
    const uint32_t t = blockIdx.x * THREADS + threadIdx.x;
    if( t >= triCount )
    {
        return;
    }
    const GeomRecord& g = geoms[ triGeom[ t ] ];
    if( !GeneratesNormals( g ) )
    {
        return;
    }
    const uint32_t i = t - g.firstTri;
    uint32_t       v[ 3 ];
    v3             p[ 3 ];
#pragma unroll
    for( int k = 0; k < 3; k++ )
    {
        v[ k ] = g.baseVertex + ( g.baseIndex == 0xFFFFFFFFu ? 3 * i + k : indices[ g.baseIndex + 3 * i + k ] );
        p[ k ] = mk3( verts[ v[ k ] ].position[ 0 ], verts[ v[ k ] ].position[ 1 ], verts[ v[ k ] ].position[ 2 ] );
    }
    const v3 n = normalize( cross( p[ 1 ] - p[ 0 ], p[ 2 ] - p[ 0 ] ) );
#pragma unroll
    for( int k = 0; k < 3; k++ )
    {
        if( g.baseIndex == 0xFFFFFFFFu || lastTri[ v[ k ] ] == i + 1u )
        {
            verts[ v[ k ] ].normal[ 0 ] = n.x;
            verts[ v[ k ] ].normal[ 1 ] = n.y;
            verts[ v[ k ] ].normal[ 2 ] = n.z;
        }
    }
}

__global__ __launch_bounds__( THREADS ) void k_compare( const uint32_t* a, const uint32_t* b, size_t words, uint32_t* diff )
{
    bool d = false;
    for( size_t i = size_t( blockIdx.x ) * THREADS + threadIdx.x; i < words; i += size_t( gridDim.x ) * THREADS )
    {
        d |= a[ i ] != b[ i ];
    }
    if( __ballot( d ) != 0ull && ( threadIdx.x & 63 ) == 0 )
    {
        atomicOr( diff, 1u );
    }
}
} // namespace

void UploadScene( SceneCollector& sc, DeviceScene& ds, hipStream_t s )
{
    sc.Seal();
    const uint32_t n = sc.triCount;
    ds.triCount       = n;
    std::memcpy( ds.classTris, sc.classTris, sizeof( ds.classTris ) );
    ds.reflRefrGeometry = false;
    ds.alphaTested      = false;
    ds.generateNormals  = false;
    for( size_t i = 0; i < sc.geoms.size; i++ )
    {
        ds.generateNormals = ds.generateNormals || ( sc.geoms.data[ i ].flags & RG_MESH_PRIMITIVE_DONT_GENERATE_NORMALS ) == 0u;
        ds.alphaTested = ds.alphaTested || ( ( sc.geoms.data[ i ].flags & RG_MESH_PRIMITIVE_ALPHA_TESTED ) && sc.geoms.data[ i ].texture != TEX_NONE );
        constexpr uint32_t RR = RG_MESH_PRIMITIVE_MIRROR | RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_GLASS | RG_MESH_PRIMITIVE_ACID |
                                RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH | RG_MESH_PRIMITIVE_MIRROR_IF_SMOOTH;
        ds.reflRefrGeometry = ds.reflRefrGeometry || ( sc.geoms.data[ i ].flags & RR ) != 0u;
    }
    ds.geomCount     = uint32_t( sc.geoms.size );
    if( n == 0 )
    {
        return;
    }
    ds.vertexCount = uint32_t( sc.vertices.size );
    ds.vertices.Reserve( sc.vertices.size * sizeof( RgPrimitiveVertex ) );
    ds.indices.Reserve( ( sc.indices.size + 1 ) * 4 );
    ds.geoms.Reserve( sc.geoms.size * sizeof( GeomRecord ) );
    RG_HIP_CHECK( hipMemcpyAsync( ds.vertices.ptr, sc.vertices.data, sc.vertices.size * sizeof( RgPrimitiveVertex ), hipMemcpyHostToDevice, s ) );
    if( sc.indices.size )
    {
        RG_HIP_CHECK( hipMemcpyAsync( ds.indices.ptr, sc.indices.data, sc.indices.size * 4, hipMemcpyHostToDevice, s ) );
    }
    RG_HIP_CHECK( hipMemcpyAsync( ds.geoms.ptr, sc.geoms.data, sc.geoms.size * sizeof( GeomRecord ), hipMemcpyHostToDevice, s ) );
}

bool AssembleScene( DeviceScene& ds, hipStream_t s )
{
    const uint32_t n = ds.triCount;
    if( n == 0 )
    {
        ds.bvhValid = false;
        return false;
    }
    // Keep the previous build's triangles to detect an unchanged scene.
    const bool canCompare = ds.bvhValid && ds.builtTriCount == n;
    if( canCompare )
    {
        std::swap( ds.tris.ptr, ds.trisPrev.ptr );
        std::swap( ds.tris.bytes, ds.trisPrev.bytes );
    }
    ds.tris.Reserve( size_t( n ) * 36 );
    ds.triNormals.Reserve( size_t( n ) * 36 );
    ds.triGeom.Reserve( size_t( n ) * 4 );
    ds.diffFlag.Reserve( 4 );

    k_tri_geom<<< DivUp( n, THREADS ), THREADS, 0, s >>>( ds.geoms.as< GeomRecord >(), ds.geomCount, n, ds.triGeom.as< uint32_t >() );
    CheckLaunch( "k_tri_geom" );
    if( ds.generateNormals )
    {
        // vertex preprocessing before the acceleration structure, as
        // Scene::SubmitForFrame does (Source/Scene.cpp:99-100)
        ds.lastTri.Reserve( size_t( ds.vertexCount ) * 4 + 4 );
        RG_HIP_CHECK( hipMemsetAsync( ds.lastTri.ptr, 0, size_t( ds.vertexCount ) * 4, s ) );
        k_flat_last<<< DivUp( n, THREADS ), THREADS, 0, s >>>( ds.indices.as< uint32_t >(), ds.geoms.as< GeomRecord >(), ds.triGeom.as< uint32_t >(), n,
                                                               ds.lastTri.as< uint32_t >() );
        CheckLaunch( "k_flat_last" );
        k_flat_write<<< DivUp( n, THREADS ), THREADS, 0, s >>>( ds.vertices.as< RgPrimitiveVertex >(), ds.indices.as< uint32_t >(),
                                                                ds.geoms.as< GeomRecord >(), ds.triGeom.as< uint32_t >(), n, ds.lastTri.as< uint32_t >() );
        CheckLaunch( "k_flat_write" );
    }
    if( ds.alphaTested )
    {
        ds.alphaRecs.Reserve( size_t( n ) * sizeof( AlphaRec ) );
    }
    k_assemble<<< DivUp( n, THREADS ), THREADS, 0, s >>>( ds.vertices.as< RgPrimitiveVertex >(), ds.indices.as< uint32_t >(), ds.geoms.as< GeomRecord >(),
                                                          ds.triGeom.as< uint32_t >(), n, ds.tris.as< float >(), ds.triNormals.as< float >(),
                                                          ds.alphaTested ? ds.alphaRecs.as< AlphaRec >() : nullptr );
    CheckLaunch( "k_assemble" );

    if( !canCompare )
    {
        return true;
    }
    RG_HIP_CHECK( hipMemsetAsync( ds.diffFlag.ptr, 0, 4, s ) );
    k_compare<<< 1024, THREADS, 0, s >>>( ds.tris.as< uint32_t >(), ds.trisPrev.as< uint32_t >(), size_t( n ) * 9, ds.diffFlag.as< uint32_t >() );
    CheckLaunch( "k_compare" );
    uint32_t diff = 1;
    RG_HIP_CHECK( hipMemcpyAsync( &diff, ds.diffFlag.ptr, 4, hipMemcpyDeviceToHost, s ) );
    RG_HIP_CHECK( hipStreamSynchronize( s ) );
    if( diff == 0 )
    {
        // identical triangles: the existing BVH (built from the same bits) stays valid
        std::swap( ds.tris.ptr, ds.trisPrev.ptr );
        std::swap( ds.tris.bytes, ds.trisPrev.bytes );
        return false;
    }
    return true;
}

} // namespace rg
