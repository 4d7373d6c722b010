// Per-frame geometry and lights of an instance, as collected by
// rgUploadMeshPrimitive / rgUpload*Light and assembled on the device.
#pragma once

#include "rg_internal.h"
#include "rg_textures.h"

#include <array>
#include <unordered_map>
#include <vector>

namespace rg
{

// One uploaded primitive (the reference's ShGeometryInstance,
// Source/Generated/ShaderCommonC.h:306-342, reduced to what the path reads).
struct GeomRecord
{
    float    model[ 12 ];       // row-major 3x4 (RgTransform)
    float    prevFromCur[ 12 ]; // row-major 3x4: prevModel * model^-1 (identity if new)
    uint32_t baseVertex;
    uint32_t baseIndex;         // 0xFFFFFFFF: non-indexed
    uint32_t firstTri;
    uint32_t triCount;
    uint32_t color;             // RgColor4DPacked32, albedo layer 0 (HitInfo.inl:126-127)
    float    emissive;          // saturate(emissive) (VertexCollector.cpp:328)
    float    roughness;         // default 1 (VertexCollector.cpp:325)
    float    metallic;          // default 0 (VertexCollector.cpp:326)
    uint32_t flags;             // RgMeshPrimitiveFlags
    uint32_t isStatic;          // prevFromCur is the identity
    uint32_t texture;           // albedo texture id (pTextureName), TEX_NONE if none
    uint32_t texOrm;            // pTextureName + "_orm": occlusion-roughness-metallic
    uint32_t texNormal;         // pTextureName + "_n": tangent-space normal
    uint32_t texEmissive;       // pTextureName + "_e": emission
    uint32_t pad[ 2 ];
};
static_assert( sizeof( GeomRecord ) == 160, "GeomRecord" );

// ShLightEncoded (Source/Generated/ShaderCommonC.h:350-357), 64 bytes.
struct LightEncoded
{
    float    color[ 3 ];
    uint32_t lightType;
    float    data0[ 4 ];
    float    data1[ 4 ];
    float    data2[ 4 ];
};
static_assert( sizeof( LightEncoded ) == 64, "LightEncoded" );

constexpr uint32_t LIGHT_TYPE_NONE        = 0;
constexpr uint32_t LIGHT_TYPE_DIRECTIONAL = 1;
constexpr uint32_t LIGHT_TYPE_SPHERE      = 2;
constexpr uint32_t LIGHT_TYPE_TRIANGLE    = 3;
constexpr uint32_t LIGHT_TYPE_SPOT        = 4;
constexpr uint32_t LIGHT_INDEX_NONE       = 32767;

// Pinned, growable host array for staging uploads.
template< typename T >
struct PinnedArray
{
    T*     data     = nullptr;
    size_t size     = 0;
    size_t capacity = 0;

    PinnedArray() = default;
    PinnedArray( const PinnedArray& ) = delete;
    PinnedArray& operator=( const PinnedArray& ) = delete;
    ~PinnedArray()
    {
        if( data )
        {
            ( void )hipHostFree( data );
        }
    }
    void Clear() { size = 0; }
    // Returns the pinned memory to the system.
    void Free()
    {
        if( data )
        {
            ( void )hipHostFree( data );
        }
        data     = nullptr;
        size     = 0;
        capacity = 0;
    }
    void Reserve( size_t n )
    {
        if( n <= capacity )
        {
            return;
        }
        size_t cap = capacity ? capacity : 1024;
        while( cap < n )
        {
            cap *= 2;
        }
        T* nd = nullptr;
        RG_HIP_CHECK( hipHostMalloc( reinterpret_cast< void** >( &nd ), cap * sizeof( T ), hipHostMallocDefault ) );
        if( data )
        {
            std::memcpy( nd, data, size * sizeof( T ) );
            ( void )hipHostFree( data );
        }
        data     = nd;
        capacity = cap;
    }
    T* Append( size_t n )
    {
        Reserve( size + n );
        T* p = data + size;
        size += n;
        return p;
    }
};

struct SceneCollector
{
    PinnedArray< RgPrimitiveVertex > vertices;
    PinnedArray< uint32_t >          indices;
    PinnedArray< GeomRecord >        geoms;
    uint32_t                         triCount = 0;
    // Water, glass and acid primitives (the reference's PT_REFRACT BLAS,
    // VertexCollectorFilterType.cpp:215-229) collect here and follow all other
    // geometry once Seal() runs, so their triangles form one trailing range
    // that shadow and bounce rays skip.
    PinnedArray< GeomRecord >        refrGeoms;
    std::vector< uint64_t >          refrUids;
    uint32_t                         refrTriCount = 0;
    bool                             sealed       = false;
    const struct TextureManager*     textures     = nullptr; // resolves pTextureName at upload

    // uniqueObjectID << 32 | primitiveIndexInMesh -> model of the previous frame
    // (UniqueID::MakeForPrimitive, Source/UniqueID.h:29-33)
    std::unordered_map< uint64_t, std::array< float, 12 > > prevModels, curModels;
    // the same id per geometry index, this frame and the previous one
    std::vector< uint64_t > geomUids, prevGeomUids;

    // lights of this and the previous frame, matched by uniqueID
    std::vector< LightEncoded >                 lights;      // [0] directional (or NONE), then regular
    std::vector< uint64_t >                     lightIds;
    bool                                        hasDirectional = false;
    std::unordered_map< uint64_t, uint32_t >    prevLightIndex; // uniqueID -> index last frame
    std::vector< uint32_t >                     prevToCur;      // last frame's index -> this frame's
    uint32_t                                    prevLightCount = 0;
    // RgDrawFrameIlluminationParams::pLightstyleValues of the last rgDrawFrame.
    // Lights are encoded at upload, before the frame's draw call sets new
    // values, so a frame's lights use the previous draw call's styles
    // (LightManager::SetLightstyles, Source/VulkanDevice.cpp:504).
    std::vector< float >                        lightstyles;

    void BeginFrame();
    void AddPrimitive( const RgMeshInfo& mesh, const RgMeshPrimitiveInfo& prim );
    // Appends the refracting geometry after the rest (idempotent until the
    // next BeginFrame); UploadScene and GeomPrevToCur call it.
    void Seal();
    uint32_t OpaqueTriCount() const { return triCount - refrTriCount; }
    void AddDirectional( const RgDirectionalLightUploadInfo& info );
    void AddSphere( const RgSphericalLightUploadInfo& info );
    void AddSpot( const RgSpotLightUploadInfo& info );
    void AddPolygonal( const RgPolygonalLightUploadInfo& info );
    // Intensity multiplier of a light (CalculateLightStyle, Source/LightManager.cpp:366-383)
    float LightStyle( const RgLightExtraInfo& extra ) const;
    // Builds prevToCur; call once all lights of the frame are added.
    void FinishLights();
    // Last frame's geometry index -> this frame's, 0xFFFFFFFF when gone
    // (getCurrentGeometryIndexByPrev, Shaders/VertexData.inl:636-658).
    void GeomPrevToCur( std::vector< uint32_t >& out );
};

// Device side of the assembled frame geometry.
struct DeviceScene
{
    DeviceBuffer vertices, indices, geoms;
    DeviceBuffer tris;       // n * 9 floats, world space
    DeviceBuffer triNormals; // n * 9 floats, world-space vertex normals
    DeviceBuffer triGeom;    // n u32
    DeviceBuffer trisPrev;   // last built copy, to skip identical rebuilds
    DeviceBuffer diffFlag;   // 1 u32
    uint32_t     triCount = 0, geomCount = 0, builtTriCount = 0;
    uint32_t     opaqueTriCount = 0; // triangles [opaqueTriCount, triCount) refract
    uint32_t     builtOpaqueTriCount = 0;
    bool         reflRefrGeometry = false; // some geometry reflects or refracts
    bool         alphaTested      = false; // some geometry is alpha tested with a texture
    DeviceBuffer alphaRecs;                // n AlphaRec when alphaTested
    TexView      tex;                      // the instance's textures for this frame
    bool         bvhValid = false;
};

// Refracting primitives: water, glass, glass-if-smooth and acid that are not
// alpha tested (VertexCollectorFilterType.cpp:215-229).
inline bool IsRefractPrimitive( uint32_t flags )
{
    return ( flags & RG_MESH_PRIMITIVE_ALPHA_TESTED ) == 0u &&
           ( flags & ( RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_GLASS | RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH | RG_MESH_PRIMITIVE_ACID ) ) != 0u;
}

// Copies the collected geometry to the device (asynchronous).
void UploadScene( SceneCollector& sc, DeviceScene& ds, hipStream_t s );
// Assembles world-space triangles from the uploaded geometry.  Returns true
// if they differ from the triangles of the last BVH build.
bool AssembleScene( DeviceScene& ds, hipStream_t s );

} // namespace rg
