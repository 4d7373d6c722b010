/*
 * RTGL1 public C API, restated for the MI355X (gfx950) implementation.
 *
 * This header is ABI- and source-compatible with RTGL1's Include/RTGL1/RTGL1.h
 * (API version "1.03.0000", Include/RTGL1/RTGL1.h:38): every type, enumerator,
 * struct field and exported function keeps the name, order and C type of the
 * original so that an application built against RTGL1 links against
 * librtgl1_amd.so unchanged.  The layout checks at the bottom pin the sizes of
 * the structs that cross the boundary on LP64.
 *
 * Differences in behaviour (not in layout):
 *   - no window surface is required: all five surface pointers of
 *     RgInstanceCreateInfo may be null (headless rendering);
 *   - rasterised, decal, lens-flare, texture and scratch entry points are
 *     accepted and ignored (they return RG_RESULT_SUCCESS);
 *   - results are read back through RTGL1_ext.h.
 *
 * The API is MIT-licensed by its author (Sultim Tsyrendashiev); the interface
 * is reproduced for compatibility.
 */
#ifndef RTGL1_H_
#define RTGL1_H_

#include <stdint.h>

/* Calling convention and export macros are empty on Linux (RTGL1.h:26-36). */
#define RGAPI
#define RGCONV

#define RG_RTGL_VERSION_API "1.03.0000"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- basics */

#define RG_DEFINE_NON_DISPATCHABLE_HANDLE( object ) typedef struct object##_T* object;

typedef uint32_t RgBool32;
#define RG_FALSE 0
#define RG_TRUE  1

RG_DEFINE_NON_DISPATCHABLE_HANDLE( RgInstance )
#define RG_NULL_HANDLE 0

/* Result codes, same numbering as RTGL1.h:79-95. */
typedef enum RgResult
{
    RG_RESULT_SUCCESS,
    RG_RESULT_SUCCESS_FOUND_MESH,
    RG_RESULT_SUCCESS_FOUND_TEXTURE,
    RG_RESULT_WRONG_INSTANCE,
    RG_RESULT_ALREADY_INITIALIZED,
    RG_RESULT_GRAPHICS_API_ERROR,
    RG_RESULT_INTERNAL_ERROR,
    RG_RESULT_CANT_FIND_SUPPORTED_PHYSICAL_DEVICE,
    RG_RESULT_FRAME_WASNT_STARTED,
    RG_RESULT_FRAME_WASNT_ENDED,
    RG_RESULT_WRONG_FUNCTION_CALL,
    RG_RESULT_WRONG_FUNCTION_ARGUMENT,
    RG_RESULT_ERROR_CANT_FIND_HARDCODED_RESOURCES,
    RG_RESULT_ERROR_CANT_FIND_SHADER,
} RgResult;

typedef enum RgMessageSeverityFlagBits
{
    RG_MESSAGE_SEVERITY_VERBOSE = 1,
    RG_MESSAGE_SEVERITY_INFO    = 2,
    RG_MESSAGE_SEVERITY_WARNING = 4,
    RG_MESSAGE_SEVERITY_ERROR   = 8,
} RgMessageSeverityFlagBits;
typedef uint32_t RgMessageSeverityFlags;

typedef void ( *PFN_rgPrint )( const char* pMessage, RgMessageSeverityFlags flags, void* pUserData );
typedef void ( *PFN_rgOpenFile )( const char* pFilePath, void* pUserData, const void** ppOutData,
                                  uint32_t* pOutDataSize, void** ppOutFileUserHandle );
typedef void ( *PFN_rgCloseFile )( void* pFileUserHandle, void* pUserData );

/* Window-system surfaces.  Only pointers to them cross the boundary and a
 * headless instance passes null for all of them, so they stay incomplete. */
typedef struct RgWin32SurfaceCreateInfo   RgWin32SurfaceCreateInfo;
typedef struct RgMetalSurfaceCreateInfo   RgMetalSurfaceCreateInfo;
typedef struct RgWaylandSurfaceCreateInfo RgWaylandSurfaceCreateInfo;
typedef struct RgXcbSurfaceCreateInfo     RgXcbSurfaceCreateInfo;
typedef struct RgXlibSurfaceCreateInfo    RgXlibSurfaceCreateInfo;

typedef enum RgTextureSwizzling
{
    RG_TEXTURE_SWIZZLING_NULL_ROUGHNESS_METALLIC,
    RG_TEXTURE_SWIZZLING_NULL_METALLIC_ROUGHNESS,
    RG_TEXTURE_SWIZZLING_OCCLUSION_ROUGHNESS_METALLIC,
    RG_TEXTURE_SWIZZLING_OCCLUSION_METALLIC_ROUGHNESS,
    RG_TEXTURE_SWIZZLING_ROUGHNESS_METALLIC,
    RG_TEXTURE_SWIZZLING_METALLIC_ROUGHNESS,
} RgTextureSwizzling;

typedef struct RgFloat2D { float data[ 2 ]; } RgFloat2D;
typedef struct RgFloat3D { float data[ 3 ]; } RgFloat3D;
typedef struct RgFloat4D { float data[ 4 ]; } RgFloat4D;

/* ------------------------------------------------------------- instance */

/* Field order follows RTGL1.h:180-252. */
typedef struct RgInstanceCreateInfo
{
    const char*                 pAppName;
    const char*                 pAppGUID;

    /* The reference requires exactly one non-null surface; here all may be null. */
    RgWin32SurfaceCreateInfo*   pWin32SurfaceInfo;
    RgMetalSurfaceCreateInfo*   pMetalSurfaceCreateInfo;
    RgWaylandSurfaceCreateInfo* pWaylandSurfaceCreateInfo;
    RgXcbSurfaceCreateInfo*     pXcbSurfaceCreateInfo;
    RgXlibSurfaceCreateInfo*    pXlibSurfaceCreateInfo;

    const char*                 pConfigPath;
    const char*                 pOverrideFolderPath;

    PFN_rgPrint                 pfnPrint;
    void*                       pUserPrintData;
    RgMessageSeverityFlags      allowedMessages;

    uint32_t                    primaryRaysMaxAlbedoLayers;
    uint32_t                    indirectIlluminationMaxAlbedoLayers;

    /* If true, front-facing (counter-clockwise) triangles are culled for all rays. */
    RgBool32                    rayCullBackFacingTriangles;
    RgBool32                    allowGeometryWithSkyFlag;
    RgBool32                    allowTexCoordLayer1;
    RgBool32                    allowTexCoordLayer2;
    RgBool32                    allowTexCoordLayer3;
    uint32_t                    lightmapTexCoordLayerIndex;

    uint32_t                    rasterizedMaxVertexCount;
    uint32_t                    rasterizedMaxIndexCount;
    RgBool32                    rasterizedVertexColorGamma;
    uint32_t                    rasterizedSkyCubemapSize;

    RgBool32                    textureSamplerForceMinificationFilterLinear;
    RgBool32                    textureSamplerForceNormalMapFilterLinear;

    RgTextureSwizzling          pbrTextureSwizzling;

    RgBool32                    effectWipeIsUsed;

    RgFloat3D                   worldUp;
    RgFloat3D                   worldForward;
    float                       worldScale;
} RgInstanceCreateInfo;

RGAPI RgResult RGCONV rgCreateInstance( const RgInstanceCreateInfo* pInfo, RgInstance* pResult );
RGAPI RgResult RGCONV rgDestroyInstance( RgInstance instance );

/* ------------------------------------------------------------- geometry */

/* Row-major 3x4 affine transform (RTGL1.h:258-262). */
typedef struct RgTransform { float matrix[ 3 ][ 4 ]; } RgTransform;
typedef struct RgMatrix3D  { float matrix[ 3 ][ 3 ]; } RgMatrix3D;

typedef struct RgExtent2D
{
    uint32_t width;
    uint32_t height;
} RgExtent2D;

typedef struct RgViewport
{
    float x;
    float y;
    float width;
    float height;
    float minDepth;
    float maxDepth;
} RgViewport;

typedef uint32_t RgColor4DPacked32;

typedef enum RgMeshPrimitiveFlagBits
{
    RG_MESH_PRIMITIVE_ALPHA_TESTED          = 1,
    RG_MESH_PRIMITIVE_TRANSLUCENT           = 2,
    RG_MESH_PRIMITIVE_FIRST_PERSON          = 4,
    RG_MESH_PRIMITIVE_FIRST_PERSON_VIEWER   = 8,
    RG_MESH_PRIMITIVE_SKY                   = 16,
    RG_MESH_PRIMITIVE_MIRROR                = 32,
    RG_MESH_PRIMITIVE_GLASS                 = 64,
    RG_MESH_PRIMITIVE_WATER                 = 128,
    RG_MESH_PRIMITIVE_DONT_GENERATE_NORMALS = 256,
    RG_MESH_PRIMITIVE_FORCE_EXACT_NORMALS   = 512,
    RG_MESH_PRIMITIVE_MIRROR_IF_SMOOTH      = 1024,
    RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH       = 2048,
    RG_MESH_PRIMITIVE_IGNORE_REFRACT_AFTER  = 4096,
    RG_MESH_PRIMITIVE_ACID                  = 8192,
    RG_MESH_PRIMITIVE_THIN_MEDIA            = 16384,
    RG_MESH_PRIMITIVE_SKY_VISIBILITY        = 32768,
} RgMeshPrimitiveFlagBits;
typedef uint32_t RgMeshPrimitiveFlags;

/* 64-byte vertex (RTGL1.h:315-322). */
typedef struct RgPrimitiveVertex
{
    float             position[ 3 ];
    uint32_t          _padding0;
    float             normal[ 3 ];
    uint32_t          _padding1;
    float             tangent[ 4 ];
    float             texCoord[ 2 ];
    RgColor4DPacked32 color;
    uint32_t          _padding2;
} RgPrimitiveVertex;

typedef struct RgEditorPortalInfo
{
    RgFloat3D inPosition;
    RgFloat3D inDirection;
    RgFloat3D outPosition;
    RgFloat3D outDirection;
} RgEditorPortalInfo;

typedef enum RgTextureLayerBlendType
{
    RG_TEXTURE_LAYER_BLEND_TYPE_OPAQUE,
    RG_TEXTURE_LAYER_BLEND_TYPE_ALPHA,
    RG_TEXTURE_LAYER_BLEND_TYPE_ADD,
    RG_TEXTURE_LAYER_BLEND_TYPE_SHADE,
} RgTextureLayerBlendType;

typedef struct RgEditorTextureLayerInfo
{
    const RgFloat2D*        pTexCoord;
    const char*             pTextureName;
    RgTextureLayerBlendType blend;
    RgColor4DPacked32       color;
} RgEditorTextureLayerInfo;

typedef struct RgEditorPBRInfo
{
    float metallicDefault;
    float roughnessDefault;
} RgEditorPBRInfo;

typedef struct RgEditorAttachedLightInfo
{
    float             intensity;
    RgColor4DPacked32 color;
} RgEditorAttachedLightInfo;

typedef struct RgEditorInfo
{
    RgBool32                  portalExists;
    RgEditorPortalInfo        portal;
    RgTextureLayerBlendType   layerBaseBlend;
    RgBool32                  layer1Exists;
    RgEditorTextureLayerInfo  layer1;
    RgBool32                  layer2Exists;
    RgEditorTextureLayerInfo  layer2;
    RgBool32                  layer3Exists;
    RgEditorTextureLayerInfo  layer3;
    RgBool32                  pbrInfoExists;
    RgEditorPBRInfo           pbrInfo;
    RgBool32                  attachedLightExists;
    RgBool32                  attachedLightEvenOnDynamic;
    RgEditorAttachedLightInfo attachedLight;
} RgEditorInfo;

/* Indexed or non-indexed triangle list with one material (RTGL1.h:385-405). */
typedef struct RgMeshPrimitiveInfo
{
    const char*              pPrimitiveNameInMesh;
    uint32_t                 primitiveIndexInMesh;
    RgMeshPrimitiveFlags     flags;

    const RgPrimitiveVertex* pVertices;
    uint32_t                 vertexCount;
    const uint32_t*          pIndices;
    uint32_t                 indexCount;

    const char*              pTextureName;
    uint32_t                 textureFrame;

    RgColor4DPacked32        color;
    float                    emissive;

    const RgEditorInfo*      pEditorInfo;
} RgMeshPrimitiveInfo;

typedef struct RgMeshInfo
{
    uint32_t    uniqueObjectID;
    const char* pMeshName;
    RgTransform transform;
    RgBool32    isExportable;
    const char* animationName;
    float       animationTime;
} RgMeshInfo;

RGAPI RgResult RGCONV rgUploadMeshPrimitive( RgInstance                 instance,
                                             const RgMeshInfo*          pMesh,
                                             const RgMeshPrimitiveInfo* pPrimitive );

RGAPI RgResult RGCONV rgUploadNonWorldPrimitive( RgInstance                 instance,
                                                 const RgMeshPrimitiveInfo* pPrimitive,
                                                 const float*               pViewProjection,
                                                 const RgViewport*          pViewport );

typedef struct RgDecalUploadInfo
{
    RgTransform transform;
    const char* pTextureName;
} RgDecalUploadInfo;

RGAPI RgResult RGCONV rgUploadDecal( RgInstance instance, const RgDecalUploadInfo* pInfo );

typedef struct RgLensFlareUploadInfo
{
    uint32_t                 vertexCount;
    const RgPrimitiveVertex* pVertices;
    uint32_t                 indexCount;
    const uint32_t*          pIndices;
    const char*              pTextureName;
    RgFloat3D                pointToCheck;
} RgLensFlareUploadInfo;

RGAPI RgResult RGCONV rgUploadLensFlare( RgInstance instance, const RgLensFlareUploadInfo* pInfo );

/* --------------------------------------------------------------- lights */

typedef struct RgLightExtraInfo
{
    RgBool32 exists;
    int      lightstyle;
    RgBool32 isVolumetric;
} RgLightExtraInfo;

/* intensity is illuminance in lux (RTGL1.h:481-493). */
typedef struct RgDirectionalLightUploadInfo
{
    uint64_t          uniqueID;
    RgBool32          isExportable;
    RgLightExtraInfo  extra;
    RgColor4DPacked32 color;
    float             intensity;
    RgFloat3D         direction;
    float             angularDiameterDegrees;
} RgDirectionalLightUploadInfo;

/* intensity is luminous flux in lumen (RTGL1.h:495-506). */
typedef struct RgSphericalLightUploadInfo
{
    uint64_t          uniqueID;
    RgBool32          isExportable;
    RgLightExtraInfo  extra;
    RgColor4DPacked32 color;
    float             intensity;
    RgFloat3D         position;
    float             radius;
} RgSphericalLightUploadInfo;

typedef struct RgPolygonalLightUploadInfo
{
    uint64_t          uniqueID;
    RgBool32          isExportable;
    RgLightExtraInfo  extra;
    RgColor4DPacked32 color;
    float             intensity;
    RgFloat3D         positions[ 3 ];
} RgPolygonalLightUploadInfo;

typedef struct RgSpotLightUploadInfo
{
    uint64_t          uniqueID;
    RgBool32          isExportable;
    RgLightExtraInfo  extra;
    RgColor4DPacked32 color;
    float             intensity;
    RgFloat3D         position;
    RgFloat3D         direction;
    float             radius;
    float             angleOuter;
    float             angleInner;
} RgSpotLightUploadInfo;

RGAPI RgResult RGCONV rgUploadDirectionalLight( RgInstance instance, const RgDirectionalLightUploadInfo* pInfo );
RGAPI RgResult RGCONV rgUploadSphericalLight( RgInstance instance, const RgSphericalLightUploadInfo* pInfo );
RGAPI RgResult RGCONV rgUploadSpotLight( RgInstance instance, const RgSpotLightUploadInfo* pInfo );
RGAPI RgResult RGCONV rgUploadPolygonalLight( RgInstance instance, const RgPolygonalLightUploadInfo* pInfo );

/* ------------------------------------------------------------- textures */

typedef enum RgSamplerFilter
{
    RG_SAMPLER_FILTER_AUTO,
    RG_SAMPLER_FILTER_LINEAR,
    RG_SAMPLER_FILTER_NEAREST,
} RgSamplerFilter;

typedef enum RgSamplerAddressMode
{
    RG_SAMPLER_ADDRESS_MODE_REPEAT,
    RG_SAMPLER_ADDRESS_MODE_CLAMP,
} RgSamplerAddressMode;

typedef struct RgOriginalTextureInfo
{
    const char*          pTextureName;
    const void*          pPixels;
    RgExtent2D           size;
    RgSamplerFilter      filter;
    RgSamplerAddressMode addressModeU;
    RgSamplerAddressMode addressModeV;
} RgOriginalTextureInfo;

typedef struct RgOriginalCubemapInfo
{
    const char* pTextureName;
    const void* pPixelsPositiveX;
    const void* pPixelsNegativeX;
    const void* pPixelsPositiveY;
    const void* pPixelsNegativeY;
    const void* pPixelsPositiveZ;
    const void* pPixelsNegativeZ;
    uint32_t    sideSize;
} RgOriginalCubemapInfo;

RGAPI RgResult RGCONV rgProvideOriginalTexture( RgInstance instance, const RgOriginalTextureInfo* pInfo );
RGAPI RgResult RGCONV rgProvideOriginalCubemapTexture( RgInstance instance, const RgOriginalCubemapInfo* pInfo );
RGAPI RgResult RGCONV rgMarkOriginalTextureAsDeleted( RgInstance instance, const char* pTextureName );

/* ---------------------------------------------------------------- frame */

typedef struct RgStartFrameInfo
{
    const char* pMapName;
    RgBool32    ignoreExternalGeometry;
} RgStartFrameInfo;

RGAPI RgResult RGCONV rgStartFrame( RgInstance instance, const RgStartFrameInfo* pInfo );

/* sType values of the RgDrawFrameInfo::pParams chain (RTGL1.h:588-601). */
typedef enum RgStructureType
{
    RG_STRUCTURE_TYPE_NONE,
    RG_STRUCTURE_TYPE_RENDER_RESOLUTION,
    RG_STRUCTURE_TYPE_ILLUMINATION,
    RG_STRUCTURE_TYPE_VOLUMETRIC,
    RG_STRUCTURE_TYPE_TONEMAPPING,
    RG_STRUCTURE_TYPE_BLOOM,
    RG_STRUCTURE_TYPE_REFLECTREFRACT,
    RG_STRUCTURE_TYPE_SKY,
    RG_STRUCTURE_TYPE_TEXTURES,
    RG_STRUCTURE_TYPE_LIGHTMAP,
    RG_STRUCTURE_TYPE_POSTEFFECTS,
} RgStructureType;

typedef enum RgSkyType
{
    RG_SKY_TYPE_COLOR,
    RG_SKY_TYPE_CUBEMAP,
    RG_SKY_TYPE_RASTERIZED_GEOMETRY,
} RgSkyType;

typedef struct RgDrawFrameTonemappingParams
{
    RgStructureType sType;
    void*           pNext;
    RgBool32        disableEyeAdaptation;
    float           ev100Min;
    float           ev100Max;
    float           luminanceWhitePoint;
    RgFloat3D       saturation;
    RgFloat3D       crosstalk;
} RgDrawFrameTonemappingParams;

typedef struct RgDrawFrameSkyParams
{
    RgStructureType sType;
    void*           pNext;
    RgSkyType       skyType;
    RgFloat3D       skyColorDefault;
    float           skyColorMultiplier;
    float           skyColorSaturation;
    RgFloat3D       skyViewerPosition;
    const char*     pSkyCubemapTextureName;
    RgMatrix3D      skyCubemapRotationTransform;
} RgDrawFrameSkyParams;

typedef struct RgDrawFrameTexturesParams
{
    RgStructureType sType;
    void*           pNext;
    RgSamplerFilter dynamicSamplerFilter;
    float           normalMapStrength;
    float           emissionMapBoost;
    float           emissionMaxScreenColor;
    float           minRoughness;
} RgDrawFrameTexturesParams;

typedef struct RgDrawFrameIlluminationParams
{
    RgStructureType sType;
    void*           pNext;
    uint32_t        maxBounceShadows;
    RgBool32        enableSecondBounceForIndirect;
    float           cellWorldSize;
    float           directDiffuseSensitivityToChange;
    float           indirectDiffuseSensitivityToChange;
    float           specularSensitivityToChange;
    float           polygonalLightSpotlightFactor;
    uint64_t*       lightUniqueIdIgnoreFirstPersonViewerShadows;
    uint32_t        lightstyleValuesCount;
    const float*    pLightstyleValues;
} RgDrawFrameIlluminationParams;

typedef struct RgDrawFrameVolumetricParams
{
    RgStructureType sType;
    void*           pNext;
    RgBool32        enable;
    float           maxHistoryLength;
    RgBool32        useSimpleDepthBased;
    float           volumetricFar;
    RgFloat3D       ambientColor;
    float           scaterring;
    float           assymetry;
    RgBool32        useIlluminationVolume;
    RgFloat3D       fallbackSourceColor;
    RgFloat3D       fallbackSourceDirection;
    float           lightMultiplier;
    RgBool32        allowTintUnderwater;
    RgFloat3D       underwaterColor;
} RgDrawFrameVolumetricParams;

typedef struct RgDrawFrameBloomParams
{
    RgStructureType sType;
    void*           pNext;
    float           bloomIntensity;
    float           inputThreshold;
    float           bloomEmissionMultiplier;
    float           lensDirtIntensity;
} RgDrawFrameBloomParams;

typedef struct RgPostEffectWipe
{
    float    stripWidth;
    RgBool32 beginNow;
    float    duration;
} RgPostEffectWipe;

typedef struct RgPostEffectRadialBlur
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
} RgPostEffectRadialBlur;

typedef struct RgPostEffectChromaticAberration
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
    float    intensity;
} RgPostEffectChromaticAberration;

typedef struct RgPostEffectInverseBlackAndWhite
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
} RgPostEffectInverseBlackAndWhite;

typedef struct RgPostEffectHueShift
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
} RgPostEffectHueShift;

typedef struct RgPostEffectDistortedSides
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
} RgPostEffectDistortedSides;

typedef struct RgPostEffectWaves
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
    float    amplitude;
    float    speed;
    float    xMultiplier;
} RgPostEffectWaves;

typedef struct RgPostEffectColorTint
{
    RgBool32  isActive;
    float     transitionDurationIn;
    float     transitionDurationOut;
    float     intensity;
    RgFloat3D color;
} RgPostEffectColorTint;

typedef struct RgPostEffectCRT
{
    RgBool32 isActive;
} RgPostEffectCRT;

typedef struct RgPostEffectTeleport
{
    RgBool32 isActive;
    float    transitionDurationIn;
    float    transitionDurationOut;
} RgPostEffectTeleport;

typedef struct RgDrawFramePostEffectsParams
{
    RgStructureType                         sType;
    void*                                   pNext;
    const RgPostEffectWipe*                 pWipe;
    const RgPostEffectRadialBlur*           pRadialBlur;
    const RgPostEffectChromaticAberration*  pChromaticAberration;
    const RgPostEffectInverseBlackAndWhite* pInverseBlackAndWhite;
    const RgPostEffectHueShift*             pHueShift;
    const RgPostEffectDistortedSides*       pDistortedSides;
    const RgPostEffectWaves*                pWaves;
    const RgPostEffectColorTint*            pColorTint;
    const RgPostEffectTeleport*             pTeleport;
    const RgPostEffectCRT*                  pCRT;
} RgDrawFramePostEffectsParams;

typedef enum RgMediaType
{
    RG_MEDIA_TYPE_VACUUM,
    RG_MEDIA_TYPE_WATER,
    RG_MEDIA_TYPE_GLASS,
    RG_MEDIA_TYPE_ACID,
} RgMediaType;

typedef struct RgDrawFrameReflectRefractParams
{
    RgStructureType sType;
    void*           pNext;
    uint32_t        maxReflectRefractDepth;
    RgMediaType     typeOfMediaAroundCamera;
    float           indexOfRefractionGlass;
    float           indexOfRefractionWater;
    float           thinMediaWidth;
    float           waterWaveSpeed;
    float           waterWaveNormalStrength;
    RgFloat3D       waterColor;
    RgFloat3D       acidColor;
    float           acidDensity;
    float           waterWaveTextureDerivativesMultiplier;
    float           waterTextureAreaScale;
    RgBool32        portalNormalTwirl;
} RgDrawFrameReflectRefractParams;

typedef enum RgRenderUpscaleTechnique
{
    RG_RENDER_UPSCALE_TECHNIQUE_LINEAR,
    RG_RENDER_UPSCALE_TECHNIQUE_NEAREST,
    RG_RENDER_UPSCALE_TECHNIQUE_AMD_FSR2,
    RG_RENDER_UPSCALE_TECHNIQUE_NVIDIA_DLSS,
} RgRenderUpscaleTechnique;

typedef enum RgRenderSharpenTechnique
{
    RG_RENDER_SHARPEN_TECHNIQUE_NONE,
    RG_RENDER_SHARPEN_TECHNIQUE_NAIVE,
    RG_RENDER_SHARPEN_TECHNIQUE_AMD_CAS,
} RgRenderSharpenTechnique;

typedef enum RgRenderResolutionMode
{
    RG_RENDER_RESOLUTION_MODE_CUSTOM,
    RG_RENDER_RESOLUTION_MODE_ULTRA_PERFORMANCE,
    RG_RENDER_RESOLUTION_MODE_PERFORMANCE,
    RG_RENDER_RESOLUTION_MODE_BALANCED,
    RG_RENDER_RESOLUTION_MODE_QUALITY,
    RG_RENDER_RESOLUTION_MODE_ULTRA_QUALITY,
} RgRenderResolutionMode;

typedef struct RgDrawFrameRenderResolutionParams
{
    RgStructureType          sType;
    void*                    pNext;
    RgRenderUpscaleTechnique upscaleTechnique;
    RgRenderSharpenTechnique sharpenTechnique;
    RgRenderResolutionMode   resolutionMode;
    RgExtent2D               customRenderSize;
    const RgExtent2D*        pPixelizedRenderSize;
    RgBool32                 resetUpscalerHistory;
} RgDrawFrameRenderResolutionParams;

typedef struct RgDrawFrameLightmapParams
{
    RgStructureType sType;
    void*           pNext;
    float           lightmapScreenCoverage;
} RgDrawFrameLightmapParams;

typedef enum RgDrawFrameRayCullFlagBits
{
    RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT = 1,
    RG_DRAW_FRAME_RAY_CULL_WORLD_1_BIT = 2,
    RG_DRAW_FRAME_RAY_CULL_WORLD_2_BIT = 4,
    RG_DRAW_FRAME_RAY_CULL_SKY_BIT     = 8,
} RgDrawFrameRayCullFlagBits;
typedef uint32_t RgDrawFrameRayCullFlags;

/* Per-frame camera and parameter chain (RTGL1.h:933-959). view is column-major. */
typedef struct RgDrawFrameInfo
{
    float                   view[ 16 ];
    float                   fovYRadians;
    float                   cameraNear;
    float                   cameraFar;
    float                   rayLength;
    RgDrawFrameRayCullFlags rayCullMaskWorld;
    RgBool32                disableRayTracedGeometry;
    RgBool32                disableRasterization;
    RgBool32                presentPrevFrame;
    double                  currentTime;
    RgBool32                vsync;
    void*                   pParams;
} RgDrawFrameInfo;

RGAPI RgResult RGCONV rgDrawFrame( RgInstance instance, const RgDrawFrameInfo* pInfo );

/* ------------------------------------------------------------ utilities */

typedef enum RgUtilImScratchTopology
{
    RG_UTIL_IM_SCRATCH_TOPOLOGY_TRIANGLES,
    RG_UTIL_IM_SCRATCH_TOPOLOGY_TRIANGLE_STRIP,
    RG_UTIL_IM_SCRATCH_TOPOLOGY_TRIANGLE_FAN,
    RG_UTIL_IM_SCRATCH_TOPOLOGY_QUADS,
} RgUtilImScratchTopology;

RGAPI RgPrimitiveVertex* RGCONV rgUtilScratchAllocForVertices( RgInstance instance, uint32_t vertexCount );
RGAPI void RGCONV rgUtilScratchFree( RgInstance instance, const RgPrimitiveVertex* pPointer );
RGAPI void RGCONV rgUtilScratchGetIndices( RgInstance              instance,
                                           RgUtilImScratchTopology topology,
                                           uint32_t                vertexCount,
                                           const uint32_t**        ppOutIndices,
                                           uint32_t*               pOutIndexCount );

RGAPI void RGCONV rgUtilImScratchClear( RgInstance instance );
RGAPI void RGCONV rgUtilImScratchStart( RgInstance instance, RgUtilImScratchTopology topology );
RGAPI void RGCONV rgUtilImScratchVertex( RgInstance instance, float x, float y, float z );
RGAPI void RGCONV rgUtilImScratchNormal( RgInstance instance, float x, float y, float z );
RGAPI void RGCONV rgUtilImScratchTexCoord( RgInstance instance, float u, float v );
RGAPI void RGCONV rgUtilImScratchTexCoord_Layer1( RgInstance instance, float u, float v );
RGAPI void RGCONV rgUtilImScratchTexCoord_Layer2( RgInstance instance, float u, float v );
RGAPI void RGCONV rgUtilImScratchTexCoord_Layer3( RgInstance instance, float u, float v );
RGAPI void RGCONV rgUtilImScratchColor( RgInstance instance, RgColor4DPacked32 color );
RGAPI void RGCONV rgUtilImScratchEnd( RgInstance instance );
RGAPI void RGCONV rgUtilImScratchSetToPrimitive( RgInstance instance, RgMeshPrimitiveInfo* pTarget );

RGAPI RgBool32 RGCONV rgUtilIsUpscaleTechniqueAvailable( RgInstance instance, RgRenderUpscaleTechnique technique );
RGAPI const char* RGCONV rgUtilGetResultDescription( RgResult result );
RGAPI RgColor4DPacked32 RGCONV rgUtilPackColorByte4D( uint8_t r, uint8_t g, uint8_t b, uint8_t a );
RGAPI RgColor4DPacked32 RGCONV rgUtilPackColorFloat4D( float r, float g, float b, float a );
RGAPI void RGCONV rgUtilExportAsTGA( RgInstance instance, const void* pPixels, uint32_t width, uint32_t height, const char* pPath );

#ifdef __cplusplus
}
#endif

/* LP64 layout pins for the structs that carry data across the boundary. */
#if defined( __cplusplus ) && defined( __LP64__ )
static_assert( sizeof( RgPrimitiveVertex ) == 64, "RgPrimitiveVertex must be 64 bytes" );
static_assert( sizeof( RgTransform ) == 48, "RgTransform must be 48 bytes" );
static_assert( sizeof( RgMeshInfo ) == 88, "RgMeshInfo layout" );
static_assert( sizeof( RgMeshPrimitiveInfo ) == 80, "RgMeshPrimitiveInfo layout" );
static_assert( sizeof( RgDrawFrameInfo ) == 120, "RgDrawFrameInfo layout" );
static_assert( sizeof( RgDirectionalLightUploadInfo ) == 48, "RgDirectionalLightUploadInfo layout" );
static_assert( sizeof( RgSphericalLightUploadInfo ) == 48, "RgSphericalLightUploadInfo layout" );
#endif

#endif /* RTGL1_H_ */
