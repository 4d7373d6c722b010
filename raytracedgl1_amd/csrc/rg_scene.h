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

// Instance classes of a frame's geometry: the reference's TLAS instances
// (one per BLAS bucket, ASManager::SetupTLASInstanceFromBLAS,
// Source/ASManager.cpp:837-963) reduced to what decides visibility, their
// instance mask and whether rayCullMaskWorld drops them.  A frame's
// triangles are ordered by class, each class one segment of the frame BVH.
enum GeomClass : uint32_t
{
    GC_WORLD_0     = 0, // PV_WORLD_0, opaque or alpha tested
    GC_WORLD_2     = 1, // RG_MESH_PRIMITIVE_SKY_VISIBILITY (PV_WORLD_2)
    GC_REFRACT_W0  = 2, // PT_REFRACT of world 0: mask REFRACT
    GC_REFRACT_W2  = 3, // PT_REFRACT of world 2: mask REFRACT, dropped with world 2
    GC_FIRST_PERSON        = 4, // PV_FIRST_PERSON (refracting or not)
    GC_FIRST_PERSON_VIEWER = 5, // PV_FIRST_PERSON_VIEWER (refracting or not)
    GC_COUNT       = 6,
};

// Instance masks and custom-index flags (Source/Generated/ShaderCommonC.h:60-71).
constexpr uint32_t IMASK_WORLD_0 = 1u << 0, IMASK_WORLD_1 = 1u << 1, IMASK_WORLD_2 = 1u << 2;
constexpr uint32_t IMASK_REFRACT = 1u << 5, IMASK_FIRST_PERSON = 1u << 6, IMASK_FIRST_PERSON_VIEWER = 1u << 7;
constexpr uint32_t CUSTOM_DYNAMIC = 1u << 0, CUSTOM_FIRST_PERSON = 1u << 1, CUSTOM_FIRST_PERSON_VIEWER = 1u << 2, CUSTOM_SKY = 1u << 3;

// Refracting primitives: water, glass, glass-if-smooth and acid that are not
// alpha tested (VertexCollectorFilterType.cpp:215-229).
__host__ __device__ inline bool IsRefractPrimitive( uint32_t flags )
{
    return ( flags & RG_MESH_PRIMITIVE_ALPHA_TESTED ) == 0u &&
           ( flags & ( RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_GLASS | RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH | RG_MESH_PRIMITIVE_ACID ) ) != 0u;
}

// VertexCollectorFilterTypeFlags_GetForGeometry (VertexCollectorFilterType.cpp:
// 233-251): first person before first-person viewer before sky visibility.
__host__ __device__ inline uint32_t GeomClassOf( uint32_t flags )
{
    if( flags & RG_MESH_PRIMITIVE_FIRST_PERSON )
    {
        return GC_FIRST_PERSON;
    }
    if( flags & RG_MESH_PRIMITIVE_FIRST_PERSON_VIEWER )
    {
        return GC_FIRST_PERSON_VIEWER;
    }
    const bool w2 = ( flags & RG_MESH_PRIMITIVE_SKY_VISIBILITY ) != 0u;
    if( IsRefractPrimitive( flags ) )
    {
        return w2 ? GC_REFRACT_W2 : GC_REFRACT_W0;
    }
    return w2 ? GC_WORLD_2 : GC_WORLD_0;
}

// The class's instance mask: first-person masks stay as they are, world
// refracting instances are rewritten to REFRACT (ASManager.cpp:866-933).
__host__ __device__ inline uint32_t GeomClassMask( uint32_t c )
{
    switch( c )
    {
        case GC_WORLD_0: return IMASK_WORLD_0;
        case GC_WORLD_2: return IMASK_WORLD_2;
        case GC_REFRACT_W0:
        case GC_REFRACT_W2: return IMASK_REFRACT;
        case GC_FIRST_PERSON: return IMASK_FIRST_PERSON;
        default: return IMASK_FIRST_PERSON_VIEWER;
    }
}

// Whether the class is in the frame's TLAS: world instances whose world bit
// is missing from rayCullMaskWorld are left out (ASManager.cpp:880-918).
__host__ __device__ inline bool GeomClassPresent( uint32_t c, uint32_t worldMask )
{
    switch( c )
    {
        case GC_WORLD_0:
        case GC_REFRACT_W0: return ( worldMask & IMASK_WORLD_0 ) != 0u;
        case GC_WORLD_2:
        case GC_REFRACT_W2: return ( worldMask & IMASK_WORLD_2 ) != 0u;
        default: return true;
    }
}

// instanceCustomIndex of a primitive's instance (ASManager.cpp:859-912): every
// uploaded primitive is dynamic; world-2 instances carry SKY when the
// instance allows geometry with the sky flag.
__host__ __device__ inline uint32_t CustomIndexOf( uint32_t flags, bool skyGeometry )
{
    const uint32_t c = GeomClassOf( flags );
    uint32_t       r = CUSTOM_DYNAMIC;
    if( c == GC_FIRST_PERSON )
    {
        r |= CUSTOM_FIRST_PERSON;
    }
    else if( c == GC_FIRST_PERSON_VIEWER )
    {
        r |= CUSTOM_FIRST_PERSON_VIEWER;
    }
    else if( skyGeometry && ( c == GC_WORLD_2 || c == GC_REFRACT_W2 ) )
    {
        r |= CUSTOM_SKY;
    }
    return r;
}

struct SceneCollector
{
    PinnedArray< RgPrimitiveVertex > vertices;
    PinnedArray< uint32_t >          indices;
    PinnedArray< GeomRecord >        geoms;
    uint32_t                         triCount = 0;
    // Geometry of world 0 (the common class) collects in geoms directly,
    // the other classes here; Seal() appends them in class order, so each
    // class is one trailing range of triangles.
    std::vector< GeomRecord >        classGeoms[ GC_COUNT ];
    std::vector< uint64_t >          classUids[ GC_COUNT ];
    uint32_t                         classTris[ GC_COUNT ] = {};
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
    void AddDirectional( const RgDirectionalLightUploadInfo& info );
    void AddSphere( const RgSphericalLightUploadInfo& info );
    void AddSpot( const RgSpotLightUploadInfo& info );
    void AddPolygonal( const RgPolygonalLightUploadInfo& info );
    // Intensity multiplier of a light (CalculateLightStyle, Source/LightManager.cpp:366-383)
    float LightStyle( const RgLightExtraInfo& extra ) const;
    // Builds prevToCur; call once all lights of the frame are added.
    void FinishLights();
    // This frame's array index of the light with a uniqueID, LIGHT_INDEX_NONE
    // if none or no id (LightManager::GetLightIndexForShaders,
    // Source/LightManager.cpp:673-689).
    uint32_t LightIndexOf( const uint64_t* uniqueId ) const;
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
    DeviceBuffer lastTri;    // per vertex: 1 + the last triangle of its geometry that uses it (flat normals)
    uint32_t     triCount = 0, geomCount = 0, builtTriCount = 0, vertexCount = 0;
    uint32_t     classTris[ GC_COUNT ] = {};      // triangles per GeomClass, in class order
    uint32_t     builtClassTris[ GC_COUNT ] = {}; // the same for the last BVH build
    bool         generateNormals = false;         // some geometry generates flat normals
    bool         reflRefrGeometry = false; // some geometry reflects or refracts
    bool         alphaTested      = false; // some geometry is alpha tested with a texture
    DeviceBuffer alphaRecs;                // n AlphaRec when alphaTested
    TexView      tex;                      // the instance's textures for this frame
    bool         bvhValid = false;
};

// Copies the collected geometry to the device (asynchronous).
void UploadScene( SceneCollector& sc, DeviceScene& ds, hipStream_t s );
// Assembles world-space triangles from the uploaded geometry.  Returns true
// if they differ from the triangles of the last BVH build.
bool AssembleScene( DeviceScene& ds, hipStream_t s );

} // namespace rg
