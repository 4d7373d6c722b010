/*
 * RTGL1 extensions of the MI355X implementation.
 *
 * RTGL1 has no readback and no stage-level entry points: its passes run
 * inside rgDrawFrame (Source/VulkanDevice.cpp:484-764) and results only reach
 * the swapchain.  These functions expose each hot-path stage on device
 * buffers so that an application (or a test harness) can inject inputs,
 * run one stage and read its output.  Every function returns RgResult and
 * never throws; a stream argument is a hipStream_t passed as void* (null =
 * the default stream).  All pointers marked d_ are device pointers.
 *
 * The reference interface each entry point stands in for is cited beside it.
 */
#ifndef RTGL1_EXT_H_
#define RTGL1_EXT_H_

#include <stddef.h>
#include <stdint.h>

#include "RTGL1.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------- device memory */
/* Plain HIP memory helpers so that callers without their own HIP runtime
 * (ctypes, cgo, JNI) can stage inputs.  No counterpart in the reference
 * (its staging is VMA-mapped memory, Source/MemoryAllocator.cpp). */
RGAPI RgResult RGCONV rgExtDeviceAlloc( size_t bytes, void** pOut );
RGAPI RgResult RGCONV rgExtDeviceFree( void* d_ptr );
/* The blocking copies and the memset first wait for all work on the device
 * (an instance keeps up to two frames in flight on its own stream). */
RGAPI RgResult RGCONV rgExtCopyHostToDevice( void* d_dst, const void* src, size_t bytes );
RGAPI RgResult RGCONV rgExtCopyDeviceToHost( void* dst, const void* d_src, size_t bytes );
RGAPI RgResult RGCONV rgExtMemsetDevice( void* d_dst, int value, size_t bytes );
/* Device-to-device copy by a plain kernel (k_copy16) on `stream`: every byte
 * read once and written once.  Pointers and size 16-byte aligned. */
RGAPI RgResult RGCONV rgExtCopyDeviceToDevice( void* d_dst, const void* d_src, size_t bytes, void* stream );
RGAPI RgResult RGCONV rgExtSynchronize( void );
/* Number of visible HIP devices (0 without a GPU). */
RGAPI int RGCONV rgExtGetDeviceCount( void );
/* Last error message of this thread ("" if none). */
RGAPI const char* RGCONV rgExtGetLastErrorMessage( void );

/* --------------------------------------------------------- stage timing */
/* Elapsed GPU time between two points of a stream, measured with HIP events
 * recorded on that stream. */
typedef struct RgExtTimer_T* RgExtTimer;
RGAPI RgResult RGCONV rgExtTimerCreate( RgExtTimer* pOut );
RGAPI RgResult RGCONV rgExtTimerDestroy( RgExtTimer timer );
RGAPI RgResult RGCONV rgExtTimerStart( RgExtTimer timer, void* stream );
RGAPI RgResult RGCONV rgExtTimerStop( RgExtTimer timer, void* stream );
/* Waits for the stop event and returns start->stop in milliseconds. */
RGAPI RgResult RGCONV rgExtTimerElapsedMs( RgExtTimer timer, float* pMs );

/* ------------------------------------------------------------------ LBVH */
/* Replaces the driver acceleration-structure build behind
 * ASBuilder::BuildBottomLevel / BuildTopLevel (Source/ASBuilder.cpp:141-155,
 * 193-206; vkCmdBuildAccelerationStructuresKHR at :147,:199).
 *
 * Input: n world-space triangles, 9 floats each (v0.xyz v1.xyz v2.xyz),
 * with transforms already applied (the reference bakes them into the BLAS,
 * Source/VertexCollector.cpp:259-261).  The primitive id of a triangle is
 * its index in this array.
 *
 * Build: FP64 centroid bounds and 21-bit-per-axis quantisation to 63-bit
 * Morton codes, stable LSD radix sort of (code, index), Karras hierarchy
 * over the sorted codes (ties broken by index), bottom-up AABB refit, then
 * collapse into 8-wide nodes.  The result is deterministic: the same input
 * gives the same nodes bit for bit. */
typedef struct RgExtBvh_T* RgExtBvh;

#define RG_EXT_BVH_NODE_FP32 0u /* 256-byte nodes with FP32 child boxes     */
#define RG_EXT_BVH_NODE_FP16 1u /* 144-byte nodes: FP32 origin, child boxes as FP16
                                   offsets from it rounded outward */

typedef struct RgExtBvhBuildParams
{
    uint32_t maxLeafSize; /* triangles per leaf child, 1..8 (default 3) */
    uint32_t nodeFormat;  /* RG_EXT_BVH_NODE_* */
} RgExtBvhBuildParams;

/* Child reference encoding inside an 8-wide node. */
#define RG_EXT_BVH_CHILD_EMPTY     0xFFFFFFFFu
#define RG_EXT_BVH_CHILD_LEAF_BIT  0x80000000u
/* leaf: LEAF_BIT | (firstSortedTriangle << 3) | (count - 1) */

/* View of a built BVH: device pointers owned by the RgExtBvh object. */
typedef struct RgExtBvhView
{
    uint32_t           triangleCount;
    uint32_t           nodeCount;        /* 8-wide nodes; node 0 is the root  */
    uint32_t           nodeFormat;
    uint32_t           nodeStrideBytes;  /* 256 (FP32) or 144 (FP16)          */
    const void*        d_nodes;
    /* Sorted triangles, 3 float4 each: (v0, primId bits) (v1, 0) (v2, 0). */
    const float*       d_sortedTriangles;
    const uint64_t*    d_mortonCodes;    /* n codes, in input order           */
    const uint32_t*    d_sortedIndices;  /* n primitive ids, in Morton order  */
    /* Binary Karras tree: n-1 internal nodes; child ids < n-1 are internal,
     * child ids >= n-1 are leaves (leaf k = sorted triangle k - (n-1)).   */
    const uint32_t*    d_binaryChildren; /* 2 per internal node               */
    const float*       d_binaryBounds;   /* 6 floats per node, internal then leaves */
    double             sceneCentroidMin[ 3 ];
    double             sceneCentroidMax[ 3 ];
} RgExtBvhView;

RGAPI RgResult RGCONV rgExtBvhCreate( RgExtBvh* pOut );
RGAPI RgResult RGCONV rgExtBvhDestroy( RgExtBvh bvh );
RGAPI RgResult RGCONV rgExtBvhBuild( RgExtBvh                   bvh,
                                     const float*               d_triangles,
                                     uint32_t                   triangleCount,
                                     const RgExtBvhBuildParams* pParams,
                                     void*                      stream );
RGAPI RgResult RGCONV rgExtBvhGetView( RgExtBvh bvh, RgExtBvhView* pOut );

/* ------------------------------------------------------------- traversal */
/* Replaces traceRayEXT (Source/Shaders/RaygenCommon.h:223,241,259,338):
 * closest hit with [tmin, tmax], or any hit for shadow rays
 * (gl_RayFlagsSkipClosestHitShaderEXT, RaygenCommon.h:337-344). */
typedef struct RgExtRay
{
    float origin[ 3 ];
    float tmin;
    float direction[ 3 ];
    float tmax;
} RgExtRay;

/* u, v: barycentric weights of v1 and v2 (hit = (1-u-v) v0 + u v1 + v v2,
 * as Source/Shaders/HitInfo.inl:239-240).  primitiveId = RG_EXT_MISS and
 * t = tmax on a miss.  For any-hit rays only primitiveId is meaningful. */
typedef struct RgExtHit
{
    float    t;
    float    u;
    float    v;
    uint32_t primitiveId;
} RgExtHit;

#define RG_EXT_MISS 0xFFFFFFFFu

#define RG_EXT_TRACE_CLOSEST         0u
#define RG_EXT_TRACE_ANY_HIT         1u /* stop at the first hit (occlusion) */
#define RG_EXT_TRACE_CULL_FRONT_FACE 2u /* gl_RayFlagsCullFrontFacingTrianglesEXT, CCW = front */
#define RG_EXT_TRACE_COLLECT_STATS   4u /* count node visits and triangle tests */

RGAPI RgResult RGCONV rgExtTraceRays( RgExtBvh        bvh,
                                      const RgExtRay* d_rays,
                                      uint32_t        rayCount,
                                      RgExtHit*       d_hits,
                                      uint32_t        flags,
                                      void*           stream );
/* Allocation-free rgExtTraceRays, for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  No counterpart in the reference.  Same arguments and
 * results as rgExtTraceRays on a single-scene BVH (rgExtBvhBuild).  Its
 * scratch is one fixed device array per GPU of 40 MiB + 8.25 KiB, held
 * until the process ends: the ray-chunk counters, the error word, the
 * statistics and the global traversal stacks of a persistent grid of at
 * most 1024 blocks (64 stack entries per ray, 24 of them in LDS; a deeper
 * stack sets RG_EXT_TRACE_ERROR_STACK_OVERFLOW).  Every call on a GPU
 * shares that array, so calls that overlap on one GPU race.  After a call,
 * rgExtBvhGetTraceErrors / rgExtBvhGetTraceStats of the BVH report that
 * call's error bits and counts.  More than RG_EXT_TRACE_NOALLOC_MAX_RAYS
 * rays are rejected with RG_RESULT_WRONG_FUNCTION_ARGUMENT. */
#define RG_EXT_TRACE_NOALLOC_MAX_RAYS 0xF0000000u
RGAPI RgResult RGCONV rgExtTraceRaysNoAlloc( RgExtBvh        bvh,
                                             const RgExtRay* d_rays,
                                             uint32_t        rayCount,
                                             RgExtHit*       d_hits,
                                             uint32_t        flags,
                                             void*           stream );
/* ------------------------------------------------------ batched scenes */
/* Several scenes in one BVH object, built by one pass sequence whose every
 * kernel covers all scenes: the triangles of scene s are
 * [sceneTriangleOffsets[s], sceneTriangleOffsets[s+1]) of d_triangles (host
 * array, sceneCount + 1 entries, each scene >= 2 triangles, 1..128 scenes).
 * Each scene is quantised in its own centroid bounds; the sort key holds the
 * scene in its top 7 bits above 19 Morton bits per axis, so the Karras
 * hierarchy splits at scene boundaries first and every scene has its own
 * root.  No counterpart in the reference, which builds one BLAS per
 * geometry bucket through the driver (ASBuilder::BuildBottomLevel,
 * Source/ASBuilder.cpp:95-155). */
RGAPI RgResult RGCONV rgExtBvhBuildBatch( RgExtBvh                   bvh,
                                          const float*               d_triangles,
                                          const uint32_t*            sceneTriangleOffsets,
                                          uint32_t                   sceneCount,
                                          const RgExtBvhBuildParams* pParams,
                                          void*                      stream );
/* Root node index of every scene of a batched BVH (sceneCount entries). */
RGAPI RgResult RGCONV rgExtBvhGetSceneRoots( RgExtBvh bvh, uint32_t* pRoots, uint32_t capacity, uint32_t* pSceneCount );
/* One persistent launch over the rays of all scenes of a batched BVH: rays
 * of scene s are [rayOffsets[s], rayOffsets[s+1]) of d_rays (host array,
 * sceneCount + 1 entries); hits report primitive ids within the scene. */
RGAPI RgResult RGCONV rgExtTraceRaysBatch( RgExtBvh        bvh,
                                           const RgExtRay* d_rays,
                                           const uint32_t* rayOffsets,
                                           RgExtHit*       d_hits,
                                           uint32_t        flags,
                                           void*           stream );

/* Traversal tuning chosen by rgExtTraceAutotune. */
typedef struct RgExtTraceTuning
{
    uint32_t stackLdsEntries; /* LDS part of the 128-entry per-lane stack: 16, 24, 32 */
    uint32_t rayChunk;        /* rays a wave takes per counter atomic: 32, 64, 128   */
    float    msPerLaunch;     /* best time on the tuning rays                         */
    float    msDefault;       /* time with the defaults (24, 64)                      */
} RgExtTraceTuning;

/* Times every stack depth x ray chunk on the given rays (rayOffsets for a
 * batched BVH, else rayCount) and keeps the fastest for this BVH object's
 * later traces; results do not depend on the choice.  Synchronises the
 * stream.  No counterpart in the reference (the driver schedules its own
 * traversal). */
RGAPI RgResult RGCONV rgExtTraceAutotune( RgExtBvh          bvh,
                                          const RgExtRay*   d_rays,
                                          uint32_t          rayCount,
                                          const uint32_t*   rayOffsets,
                                          RgExtHit*         d_hits,
                                          uint32_t          flags,
                                          void*             stream,
                                          RgExtTraceTuning* pOut );

/* Node visits and triangle tests of the last rgExtTraceRays call made with
 * RG_EXT_TRACE_COLLECT_STATS (synchronises the device). */
RGAPI RgResult RGCONV rgExtBvhGetTraceStats( RgExtBvh bvh, uint64_t* pNodeVisits, uint64_t* pTriangleTests );
/* Error bits of the last rgExtTraceRays call (synchronises the device); 0 when
 * every ray was traced completely.  A set bit means some hits may be wrong. */
#define RG_EXT_TRACE_ERROR_STACK_OVERFLOW 1u /* a traversal stack entry was dropped */
#define RG_EXT_TRACE_ERROR_STEP_LIMIT     2u /* a ray exceeded the step limit (malformed tree) */
RGAPI RgResult RGCONV rgExtBvhGetTraceErrors( RgExtBvh bvh, uint32_t* pErrors );

/* -------------------------------------------------------------- tonemap */
/* CmLuminanceHistogram.comp:41-86: 256-bin log2-luminance histogram of a
 * width x height FP32 luminance image, accumulated (added) into d_histogram. */
RGAPI RgResult RGCONV rgExtLuminanceHistogram( const float* d_luminance,
                                               uint32_t     width,
                                               uint32_t     height,
                                               float        minLogLuminance,
                                               float        maxLogLuminance,
                                               uint32_t*    d_histogram,
                                               void*        stream );

/* CmLuminanceAvg.comp:39-90: weighted log average, exponential adaptation
 * of *d_avgLuminance (read-modify-write) and clearing of d_histogram. */
RGAPI RgResult RGCONV rgExtLuminanceAverage( uint32_t* d_histogram,
                                             uint32_t  pixelCount,
                                             float     minLogLuminance,
                                             float     maxLogLuminance,
                                             float     timeDelta,
                                             RgBool32  stopEyeAdaptation,
                                             float*    d_avgLuminance,
                                             void*     stream );

/* --------------------------------------------------------- framebuffers */
/* The per-pixel images of the hot path (Source/Generated/
 * GenerateShaderCommon.py:802-887), stored as planes in HBM.  Planes marked
 * PREV are double-buffered and swapped by rgExtFramebuffersSwap
 * (Source/Framebuffers.cpp:35-47).  Layouts: C = checkerboard W x H
 * (Source/Shaders/ShaderCommonGLSLFunc.h:306-357), R = regular W x H,
 * G = gradient strata ceil(W/3) x ceil(H/3).  Element types in the
 * FP32 storage mode are given per plane. */
typedef enum RgExtPlane
{
    RG_EXT_PLANE_NORMAL,                 /* C u32  encoded normal, PREV       */
    RG_EXT_PLANE_METALLIC_ROUGHNESS,     /* C f32x2, PREV                     */
    RG_EXT_PLANE_DEPTH_WORLD,            /* C f32 linear depth, PREV          */
    RG_EXT_PLANE_DEPTH_GRAD,             /* C f32                             */
    RG_EXT_PLANE_MOTION,                 /* C f32x4 (x, y, z, 0)              */
    RG_EXT_PLANE_UNFILTERED_DIRECT,      /* C f32x4                           */
    RG_EXT_PLANE_UNFILTERED_SPECULAR,    /* C f32x4                           */
    RG_EXT_PLANE_UNFILTERED_INDIR,       /* C f32x4                           */
    RG_EXT_PLANE_SURFACE_POSITION,       /* C f32x4, PREV                     */
    RG_EXT_PLANE_VISIBILITY_BUFFER,      /* C u32x4, PREV                     */
    RG_EXT_PLANE_VIEW_DIRECTION,         /* C f32x4 (dir, hit distance)       */
    RG_EXT_PLANE_THROUGHPUT,             /* C f32x4                           */
    RG_EXT_PLANE_IS_SKY,                 /* C u32                             */
    RG_EXT_PLANE_ACCUM_HISTORY_LENGTH,   /* C f32x4, PREV                     */
    RG_EXT_PLANE_DIFF_TEMPORARY,         /* C f32x4                           */
    RG_EXT_PLANE_DIFF_ACCUM_COLOR,       /* C f32x4                           */
    RG_EXT_PLANE_DIFF_ACCUM_MOMENTS,     /* C f32x2, PREV                     */
    RG_EXT_PLANE_DIFF_COLOR_HISTORY,     /* C f32x4 (color, variance)         */
    RG_EXT_PLANE_DIFF_PING,              /* C f32x4 (color, variance)         */
    RG_EXT_PLANE_DIFF_PONG,              /* C f32x4 (color, variance)         */
    RG_EXT_PLANE_SPEC_ACCUM_COLOR,       /* C f32x4, PREV                     */
    RG_EXT_PLANE_SPEC_PING,              /* C f32x4                           */
    RG_EXT_PLANE_SPEC_PONG,              /* C f32x4                           */
    RG_EXT_PLANE_INDIR_ACCUM,            /* C f32x4, PREV                     */
    RG_EXT_PLANE_INDIR_PING,             /* C f32x4                           */
    RG_EXT_PLANE_INDIR_PONG,             /* C f32x4                           */
    RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE, /* C f32                           */
    RG_EXT_PLANE_RESERVOIRS,             /* C u32x4, PREV                     */
    RG_EXT_PLANE_RESERVOIRS_INITIAL,     /* C u32x4                           */
    RG_EXT_PLANE_GRADIENT_INPUTS,        /* C f32x2, PREV                     */
    RG_EXT_PLANE_ALBEDO,                 /* R f32x4                           */
    RG_EXT_PLANE_HISTOGRAM_INPUT,        /* R f32                             */
    RG_EXT_PLANE_PRE_FINAL,              /* R f32x4                           */
    RG_EXT_PLANE_DIS_PING_GRADIENT,      /* G f32x4                           */
    RG_EXT_PLANE_DIS_PONG_GRADIENT,      /* G f32x4                           */
    RG_EXT_PLANE_DIS_GRADIENT_HISTORY,   /* G f32x4                           */
    RG_EXT_PLANE_GRADIENT_PREV_PIX,      /* G u32                             */
    /* ReSTIR GI (RestirBuffers.cpp:117-133; ReservoirIndirect.h:160-258),
     * the reference's 6- and 8-word records split into 16-byte-aligned planes */
    RG_EXT_PLANE_INDIR_INITIAL_SAMPLE,   /* C u32x4 (position bits xyz, encoded normal)      */
    RG_EXT_PLANE_INDIR_INITIAL_RADIANCE, /* C u32x2 (E5B9G9R9 radiance, 1/source pdf bits)   */
    RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, /* C u32x4 (position bits xyz, encoded normal), PREV */
    RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, /* C u32x4 (E5 radiance, target pdf, weight sum, M), PREV */
    RG_EXT_PLANE_FINAL,                  /* R f32x4 tone-mapped LDR colour in [0,1]            */
    /* framebufScreenEmisRT / framebufAcidFogRT (RaygenPrimary.inl:126-127,
     * :326-328, :588-589): emission seen on screen and acid fog */
    RG_EXT_PLANE_SCREEN_EMISSION,        /* R f32x4                                           */
    RG_EXT_PLANE_ACID_FOG,               /* R f32x4                                           */
    /* the denoiser's guide: decoded normal and linear depth of each pixel,
     * written by the temporal pass for the stencil passes (no counterpart in
     * the reference, whose passes decode the normal image at every tap) */
    RG_EXT_PLANE_DENOISE_GUIDE,          /* C f32x4 (normal xyz, depth)                      */
    /* what the ReSTIR GI spatial reuse reads of a neighbour, in one 64-byte
     * record written after the initial samples (no counterpart in the
     * reference, which reads the G-buffer and initial-sample images): f32x4
     * surface position xyz + linear depth, f32x4 decoded normal xyz + is-sky,
     * f32x4 initial sample position xyz + 1/source pdf, f32x4 decoded sample
     * normal xyz + E5B9G9R9 radiance code */
    RG_EXT_PLANE_GI_REUSE,               /* C u32x16                                          */
    RG_EXT_PLANE_COUNT
} RgExtPlane;

typedef struct RgExtFramebuffers_T* RgExtFramebuffers;

/* Storage of the temporal and a-trous planes (ACCUM_HISTORY_LENGTH,
 * DIFF_TEMPORARY .. INDIR_PONG, ATROUS_FILTERED_VARIANCE).  FP16 stores
 * them as IEEE half (2 bytes per component, same component count), like the
 * reference's RGBA16F / E5B9G9R9 images of these buffers
 * (GenerateShaderCommon.py:830-849); values are rounded to nearest and
 * clamped to +-65504, and all arithmetic stays FP32.  Every other plane is
 * FP32 in both modes. */
#define RG_EXT_STORAGE_FP32 0u
#define RG_EXT_STORAGE_FP16 1u

RGAPI RgResult RGCONV rgExtFramebuffersCreate( uint32_t width, uint32_t height, RgExtFramebuffers* pOut );
/* As rgExtFramebuffersCreate, with a storage mode (RG_EXT_STORAGE_*). */
RGAPI RgResult RGCONV rgExtFramebuffersCreateWithStorage( uint32_t width, uint32_t height, uint32_t storage, RgExtFramebuffers* pOut );
RGAPI RgResult RGCONV rgExtFramebuffersDestroy( RgExtFramebuffers fb );
/* Current (prev = 0) or previous-frame (prev = 1) copy of a plane.  Planes
 * without PREV return the same pointer for both. */
RGAPI RgResult RGCONV rgExtFramebuffersGetPlane( RgExtFramebuffers fb,
                                                 RgExtPlane        plane,
                                                 RgBool32          prev,
                                                 void**            d_pOut,
                                                 size_t*           pBytes );
RGAPI RgResult RGCONV rgExtFramebuffersSwap( RgExtFramebuffers fb );

/* Per-frame constants of the hot path: the fields of ShGlobalUniform
 * (Source/Generated/ShaderCommonC.h:209-304) that its kernels read.
 * Matrices are column-major like the reference's. */
typedef struct RgExtFrameParams
{
    uint32_t renderWidth;  /* must be even (Source/VulkanDevice.cpp:186-187) */
    uint32_t renderHeight;
    uint32_t frameId;
    float    timeDelta;

    float    view[ 16 ];
    float    invView[ 16 ];
    float    viewPrev[ 16 ];
    float    projection[ 16 ];
    float    invProjection[ 16 ];
    float    projectionPrev[ 16 ];
    float    cameraPosition[ 4 ];
    float    cameraPositionPrev[ 4 ];

    float    minLogLuminance;
    float    maxLogLuminance;
    uint32_t stopEyeAdaptation;
    uint32_t antiFireflyEnabled;

    float    jitterX;
    float    jitterY;
    float    primaryRayMinDist;
    float    rayLength;

    uint32_t rayCullBackFaces;
    uint32_t lightCount;          /* regular (non-directional) lights         */
    uint32_t lightCountPrev;
    uint32_t directionalLightExists;

    float    skyColorDefault[ 4 ];
    float    skyColorMultiplier;
    float    skyColorSaturation;
    float    cellWorldSize;
    uint32_t maxBounceShadowsLights;

    float    gradientMultDiffuse;
    float    gradientMultIndirect;
    float    gradientMultSpecular;
    uint32_t indirSecondBounce;

    float    minRoughness;
    float    polyLightSpotlightFactor;
    uint32_t atrousIterations;    /* 4 as the reference; 5 adds step 16       */
    uint32_t restirInitialCandidates; /* 8 as the reference                   */

    uint32_t restirTemporalSamples;   /* 1 as the reference                   */
    uint32_t restirSpatialSamples;    /* 8 as the reference                   */
    uint32_t debugShowFlags;
    float    emissionMapBoost;    /* 100 as the reference                   */

    /* Reflection / refraction chain (RgDrawFrameReflectRefractParams, filled
     * as Source/VulkanDevice.cpp:306-339 does; defaults DrawFrameInfo.h:125-146). */
    uint32_t reflectRefractMaxDepth;  /* 2; at most 16                        */
    uint32_t cameraMediaType;         /* RgMediaType around the camera       */
    float    indexOfRefractionGlass;  /* 1.52                                */
    float    indexOfRefractionWater;  /* 1.33 (also acid)                    */
    float    thinMediaWidth;          /* 0.1                                 */
    float    cameraRayConeSpreadAngle; /* atan(2 tan(fovY / 2) / renderHeight) (VulkanDevice.cpp:399-400) */
    float    normalMapStrength;       /* RgDrawFrameTexturesParams, 1          */
    float    waterWaveStrength;       /* waterWaveNormalStrength, 1            */
    float    waterColorAndDensity[ 4 ]; /* (0.3, 0.73, 0.63, 0)             */
    float    acidColorAndDensity[ 4 ];  /* (0.0, 0.66, 0.55, 10)            */
    float    worldUp[ 4 ];              /* normalised RgInstanceCreateInfo::worldUp */
    float    emissionMaxScreenColor;    /* RgDrawFrameTexturesParams, 1.5     */

    /* Ray cull masks (Source/VulkanDevice.cpp:344-393, RaygenCommon.h:103-177):
     * INSTANCE_MASK_WORLD_0/1/2 bits (1, 2, 4) of RgDrawFrameInfo::
     * rayCullMaskWorld; world instances outside it are not traced at all. */
    uint32_t rayCullMaskWorld;           /* default 7 (every world)                        */
    uint32_t rayCullMaskWorldShadow;     /* the same without world 2 when sky geometry is allowed */
    uint32_t lightIndexIgnoreFPVShadows; /* light whose shadow rays skip first-person-viewer
                                            geometry (VulkanDevice.cpp:288-289); 32767: none */
    uint32_t skyGeometry;                /* RgInstanceCreateInfo::allowGeometryWithSkyFlag:
                                            hits on world-2 geometry read as sky (RaygenCommon.h:191-205) */
    uint32_t pad2[ 3 ];
} RgExtFrameParams;

/* Fills an RgExtFrameParams with the reference defaults
 * (Source/DrawFrameInfo.h:34-214) for a width x height render. */
RGAPI RgResult RGCONV rgExtFrameParamsDefaults( uint32_t width, uint32_t height, RgExtFrameParams* pOut );

/* Primary rays of a camera into d_rays (renderWidth x renderHeight rays,
 * row-major): getRayDir over pixel centres plus jitter from invView /
 * invProjection, origin cameraPosition, [primaryRayMinDist, rayLength]
 * (RaygenPrimary.inl:300-314, ShaderCommonGLSLFunc.h:460-478). */
RGAPI RgResult RGCONV rgExtPrimaryRays( const RgExtFrameParams* pParams, RgExtRay* d_rays, void* stream );


/* A-SVGF passes, in the order of Denoiser::Denoise (Source/Denoiser.cpp:64-255). */
/* CmASVGFGradientAtrous.comp:42-105, iteration 0..3 (steps 1,2,4,8). */
RGAPI RgResult RGCONV rgExtDenoiseGradientAtrous( RgExtFramebuffers fb, const RgExtFrameParams* p, uint32_t iteration, void* stream );
/* CmSVGFTemporalAccumulation.comp:101-385. */
RGAPI RgResult RGCONV rgExtDenoiseTemporal( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream );
/* CmAntiFirefly.comp:96-159. */
RGAPI RgResult RGCONV rgExtDenoiseAntiFirefly( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream );
/* CmSVGFEstimateVariance.comp:129-206. */
RGAPI RgResult RGCONV rgExtDenoiseVariance( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream );
/* CmSVGFAtrous_Iter0.comp:196-335 (iteration 0) and CmSVGFAtrous.comp:71-349
 * (iterations 1..3; the last one also composites into PRE_FINAL and
 * HISTOGRAM_INPUT).  Iteration 4 (step 16) exists only with
 * atrousIterations = 5. */
RGAPI RgResult RGCONV rgExtDenoiseAtrous( RgExtFramebuffers fb, const RgExtFrameParams* p, uint32_t iteration, void* stream );
/* The whole chain as Denoiser::Denoise runs it. */
RGAPI RgResult RGCONV rgExtDenoise( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream );
/* No counterpart (the reference's compute shaders have fixed 16x16 groups):
 * the workgroup tile of a-trous iterations >= 1, 256 pixels of tileWidth
 * 16, 32, 64 (the default: the fastest at 1080p) or 128 columns.  Results
 * do not depend on it.
 * rgExtDenoiseAutotune times iteration 1 on the set's current planes with
 * each width (it writes that iteration's outputs and synchronises the
 * stream), keeps the fastest for later a-trous calls on the set and returns
 * it; pMs (optional) receives the four times in milliseconds, in width
 * order. */
RGAPI RgResult RGCONV rgExtFramebuffersSetAtrousTile( RgExtFramebuffers fb, uint32_t tileWidth );
RGAPI RgResult RGCONV rgExtDenoiseAutotune( RgExtFramebuffers fb, const RgExtFrameParams* p, void* stream, uint32_t* pTileWidth, float* pMs );
/* Final compose: CmCheckerboard.comp:49-76 and CmPrepareFinal.comp:148-176
 * (ImageComposition.cpp:87-156). PRE_FINAL -> FINAL: checkerboard resolve of
 * split pixels, auto exposure from *d_avgLuminance (Exposure.h:29-56) and
 * the LPM 709->709 tone map (Shaders/LPM/ffx_lpm.h, set up as
 * ImageComposition.cpp:377-394). tonemapping may be null (defaults). */
RGAPI RgResult RGCONV rgExtComposeFinal( RgExtFramebuffers framebuffers, const RgDrawFrameTonemappingParams* tonemapping,
                                         const float* d_avgLuminance, void* stream );
/* The same with the screen emission weight of RgDrawFrameTexturesParams
 * (emissionMaxScreenColor; textures may be null: 1.5):
 * FINAL = LPM(PRE_FINAL * exposure + SCREEN_EMISSION * emissionMaxScreenColor)
 * + ACID_FOG, the three resolved alike on split pixels. */
RGAPI RgResult RGCONV rgExtComposeFinalEx( RgExtFramebuffers framebuffers, const RgDrawFrameTonemappingParams* tonemapping,
                                           const RgDrawFrameTexturesParams* textures, const float* d_avgLuminance, void* stream );


/* ------------------------------------------------------ instance extras */
/* Build-only knobs, passed in the RgDrawFrameInfo::pParams chain with this
 * sType (the reference walks the chain in AccessParams, Source/
 * DrawFrameInfo.h:282-345, and logs unknown sTypes). */
#define RG_EXT_STRUCTURE_TYPE_PATH_PARAMS ( ( RgStructureType )0x52470001 )

typedef struct RgExtDrawFramePathParams
{
    RgStructureType sType;
    void*           pNext;
    uint32_t        atrousIterations;        /* 4 (reference) or 5; 0 = default */
    uint32_t        restirInitialCandidates; /* 0 = default (8)                 */
    uint32_t        restirTemporalSamples;   /* 0xFFFFFFFF = default (1)        */
    uint32_t        restirSpatialSamples;    /* 0xFFFFFFFF = default (8)        */
    uint32_t        renderWidth;             /* 0 = RENDER_RESOLUTION params / 1280 */
    uint32_t        renderHeight;            /* 0 = RENDER_RESOLUTION params / 720  */
    uint32_t        collectTraversalStats;   /* count node visits / triangle tests  */
    uint32_t        precisionFlags;          /* RG_EXT_PRECISION_* bits; 0 = FP32  */
} RgExtDrawFramePathParams;

/* FP16 precision form: FP16 storage of the denoiser's temporal and a-trous
 * planes (RG_EXT_STORAGE_FP16) and FP16 BVH nodes (RG_EXT_BVH_NODE_FP16);
 * camera matrices stay FP64 on the host and ray-triangle tests FP32. */
#define RG_EXT_PRECISION_FP16_STORAGE 1u
#define RG_EXT_PRECISION_FP16_NODES   2u

/* Per-frame statistics of the last rgDrawFrame; GPU times from HIP events
 * recorded on the instance stream around each stage.  rgDrawFrame returns
 * once its frame is enqueued (at most two frames in flight, as the
 * reference's MAX_FRAMES_IN_FLIGHT, Source/Common.h:38), so
 * rgExtGetFrameStats waits for the last frame to finish. */
typedef struct RgExtFrameStats
{
    uint32_t frameId;
    uint32_t renderWidth;
    uint32_t renderHeight;
    uint32_t triangleCount;
    uint32_t bvhRebuilt;
    uint32_t bvhNodeCount;
    uint64_t raysTraced;
    float    gpuMsTotal;
    float    gpuMsAssemble;
    float    gpuMsBvh;
    float    gpuMsTrace;
    float    gpuMsDenoise;
    float    gpuMsTonemap;
    float    avgLuminance;
    float    gpuMsUpload;       /* host-to-device copies of the uploaded geometry   */
    float    gpuMsTraversal;    /* sum over the traversal-kernel launches           */
    uint32_t traversalLaunches;
    uint32_t traversalErrors;   /* OR of RG_EXT_TRACE_ERROR_* over the frame's launches */
    uint64_t nodeVisits;        /* with collectTraversalStats, else 0               */
    uint64_t triangleTests;
} RgExtFrameStats;

RGAPI RgResult RGCONV rgExtGetFrameStats( RgInstance instance, RgExtFrameStats* pOut );
/* The instance's framebuffer set (valid after the first rgDrawFrame); owned
 * by the instance, never pass it to rgExtFramebuffersDestroy.  Waits for the
 * frames in flight, so the planes hold the last frame's results. */
RGAPI RgResult RGCONV rgExtGetFramebuffers( RgInstance instance, RgExtFramebuffers* pOut );
/* The frame constants used by the last rgDrawFrame. */
RGAPI RgResult RGCONV rgExtGetFrameParams( RgInstance instance, RgExtFrameParams* pOut );
/* The luminance histogram of the last frame (256 bins) before it is cleared. */
RGAPI RgResult RGCONV rgExtReadbackHistogram( RgInstance instance, uint32_t* pOut256 );
/* The instance's stream (hipStream_t). */
RGAPI RgResult RGCONV rgExtGetStream( RgInstance instance, void** pOut );
/* Static geometry: when enabled, a frame that uploads no primitives keeps
 * the previous frame's geometry resident in HBM (no copy, no assembly, no
 * BVH build), like the reference's static scene (ASManager::
 * SubmitStaticGeometry, Source/ASManager.cpp:696-764); uploading primitives
 * replaces it. */
RGAPI RgResult RGCONV rgExtSetStaticGeometry( RgInstance instance, RgBool32 enable );

/* ----------------------------------------------------------------- serving */
/* A render server: many camera sessions over shared static scenes, their
 * frames in flight together on a few HIP streams.  The reference renders one
 * camera per instance (VulkanDevice); a session is that instance's per-camera
 * state (A-SVGF history and moments, ReSTIR DI/GI reservoirs, light grid,
 * eye adaptation) kept resident in HBM, and a frame is rgDrawFrame's work
 * from the traced passes to the FINAL plane.
 *
 * Memory: a session owns only the planes a frame reads back from the frame
 * before (RG_EXT_PLANE_* with PREV, DIFF_COLOR_HISTORY, DIS_GRADIENT_HISTORY,
 * GRADIENT_PREV_PIX) and FINAL; the within-frame planes and wavefront
 * buffers belong to a stream and serve every session that stream runs.
 * Sessions are admitted while their bytes fit memoryBudgetBytes.
 *
 * Scheduling: a session with a frame in flight stays on that frame's
 * stream, so its frames stay in order; an idle session's next frame goes to
 * the stream with the fewest frames queued.  A dispatcher thread launches
 * waiting requests while a stream has fewer than framesInFlightPerStream
 * launches queued, and retires frames as their completion events fire.
 * The request it launches first is the oldest waiting one
 * (RG_EXT_SCHEDULE_OLDEST_FIRST) or the newest (RG_EXT_SCHEDULE_NEWEST_FIRST,
 * which keeps the latency of served frames low when more is offered than
 * the GPU can render).  A request that is still waiting when its session
 * submits again is replaced by the newer one and reported as dropped (the
 * newest camera wins); one that has waited maxQueueSeconds is dropped as
 * expired.
 *
 * Batching: with batchFrames > 1, one launch takes up to batchFrames
 * waiting requests of sessions of the same scene (and equal ray cull masks
 * and reflection depth) and renders them together: each traversal of a
 * frame stage covers every frame's rays in one persistent launch, the
 * shading, A-SVGF and compose kernels run per frame.  Each stream holds
 * batchFrames sets of within-frame planes.
 *
 * Readback: with RG_EXT_SERVER_READBACK, each frame ends with its FINAL
 * plane packed to RGBA8 (sRGB transfer, the swapchain encoding) and copied
 * to pinned host memory of the session (two images per session, used in
 * turn); a frame is finished when that copy is.  RgExtFrameResult::pImage
 * points at the frame's pixels, valid until the session finishes two more
 * frames.
 *
 * Scene groups: frames of different scenes share a launch when their scenes
 * are members of one group (rgExtSceneGroupCreate), which builds one
 * batched BVH over the members' triangles, each scene its own subtree and
 * root.  Each traversal stage of such a batch is one launch whose rays
 * start at their own scene's root.  A grouped scene's frames always trace
 * the group's BVH, alone or batched.
 *
 * A session renders the same pixels alone, among others and in a batch:
 * its frames depend only on its own requests and the scene (and, for a
 * grouped scene, the group it was traced in). */
typedef struct RgExtServer_T*     RgExtServer;
typedef struct RgExtScene_T*      RgExtScene;
typedef struct RgExtSession_T*    RgExtSession;
typedef struct RgExtSceneGroup_T* RgExtSceneGroup;

#define RG_EXT_SCHEDULE_OLDEST_FIRST 0u
#define RG_EXT_SCHEDULE_NEWEST_FIRST 1u
#define RG_EXT_SERVER_READBACK       1u

typedef struct RgExtServerCreateInfo
{
    uint32_t streamCount;             /* 1..16; 0 = 4                                  */
    uint32_t framesInFlightPerStream; /* launches queued per stream, 1..8; 0 = 2       */
    uint32_t maxRenderWidth;          /* capacity of the per-stream within-frame planes */
    uint32_t maxRenderHeight;
    uint32_t precisionFlags;          /* RG_EXT_PRECISION_* for every scene and session */
    uint32_t batchFrames;             /* frames per launch, 1..16; 0 = 1                */
    uint64_t memoryBudgetBytes;       /* session bytes admitted; 0 = no limit           */
    double   maxQueueSeconds;         /* waiting longer drops a request; 0 = no limit   */
    uint32_t schedule;                /* RG_EXT_SCHEDULE_*                              */
    uint32_t flags;                   /* RG_EXT_SERVER_READBACK                         */
} RgExtServerCreateInfo;

/* rgExtSessionCreate: the session's bytes do not fit the budget. */
#define RG_EXT_RESULT_OVER_BUDGET ( ( RgResult )0x52470101 )

#define RG_EXT_FRAME_DONE    0u
#define RG_EXT_FRAME_DROPPED 1u /* replaced by a newer request before it started */
#define RG_EXT_FRAME_FAILED  2u
#define RG_EXT_FRAME_EXPIRED 3u /* waited maxQueueSeconds without starting       */

typedef struct RgExtFrameResult
{
    RgExtSession session;
    uint64_t     tag;        /* the caller's tag from rgExtSessionSubmit            */
    uint32_t     status;     /* RG_EXT_FRAME_*                                      */
    uint32_t     frameId;    /* the session's frame index (DONE)                    */
    double       submitTime; /* seconds on the server's steady clock                */
    double       startTime;  /* the frame's work was enqueued (DONE)                */
    double       endTime;    /* completion seen (DONE; after the readback), or drop time */
    const void*  pImage;     /* RGBA8 pixels, width x height rows (DONE with READBACK) */
    uint32_t     batchSize;  /* frames in the frame's launch (DONE)                 */
    uint32_t     _pad;
} RgExtFrameResult;

typedef struct RgExtServerStats
{
    uint32_t sessions;
    uint32_t streams;
    uint64_t sessionBytes;   /* admitted against the budget                        */
    uint64_t sharedBytes;    /* per-stream within-frame planes, wavefronts, traversal stacks */
    uint64_t sceneBytes;
    uint64_t framesDone;
    uint64_t framesDropped;  /* replaced or expired                               */
    uint64_t framesFailed;
    double   now;            /* the server's steady clock, seconds                 */
} RgExtServerStats;

RGAPI RgResult RGCONV rgExtServerCreate( const RgExtServerCreateInfo* pInfo, RgExtServer* pOut );
/* Waits for every frame, then frees sessions, scenes and the server. */
RGAPI RgResult RGCONV rgExtServerDestroy( RgExtServer server );
/* A static scene: primitives and lights are collected like rgUploadMeshPrimitive
 * and rgUpload*Light, then rgExtSceneBuild uploads them and builds the BVH
 * once (synchronously). */
RGAPI RgResult RGCONV rgExtSceneCreate( RgExtServer server, RgExtScene* pOut );
RGAPI RgResult RGCONV rgExtSceneUploadMeshPrimitive( RgExtScene scene, const RgMeshInfo* pMesh, const RgMeshPrimitiveInfo* pPrimitive );
RGAPI RgResult RGCONV rgExtSceneUploadDirectionalLight( RgExtScene scene, const RgDirectionalLightUploadInfo* pInfo );
RGAPI RgResult RGCONV rgExtSceneUploadSphericalLight( RgExtScene scene, const RgSphericalLightUploadInfo* pInfo );
RGAPI RgResult RGCONV rgExtSceneBuild( RgExtScene scene );
/* Fails with RG_RESULT_WRONG_FUNCTION_CALL while sessions use the scene or
 * it is a member of a group. */
RGAPI RgResult RGCONV rgExtSceneDestroy( RgExtScene scene );
/* A group of 2..128 built scenes whose frames may share launches: every
 * primitive opaque world-0 geometry (no alpha test, no sky visibility,
 * first-person, water, glass or acid primitives), at least 2 triangles per
 * scene, the same light kinds in every scene (a directional light or not,
 * sphere lights or not).  A scene belongs to at most one group.  Builds the
 * group's BVH synchronously. */
RGAPI RgResult RGCONV rgExtSceneGroupCreate( RgExtServer server, const RgExtScene* pScenes, uint32_t sceneCount, RgExtSceneGroup* pOut );
/* Waits until no request is waiting or in flight, then frees the group; its
 * scenes trace their own BVHs again. */
RGAPI RgResult RGCONV rgExtSceneGroupDestroy( RgExtSceneGroup group );
/* A camera session of a fixed render size over a built scene; returns
 * RG_EXT_RESULT_OVER_BUDGET when its bytes do not fit. */
RGAPI RgResult RGCONV rgExtSessionCreate( RgExtScene scene, uint32_t width, uint32_t height, RgExtSession* pOut );
/* Drops its waiting request, waits for its frames in flight, frees it. */
RGAPI RgResult RGCONV rgExtSessionDestroy( RgExtSession session );
/* Requests a frame: camera, time and pParams as for rgDrawFrame (the render
 * size is the session's; RENDER_RESOLUTION is ignored).  The parameters are
 * copied; the call returns at once. */
RGAPI RgResult RGCONV rgExtSessionSubmit( RgExtSession session, const RgDrawFrameInfo* pInfo, uint64_t tag );
/* Results of finished and dropped requests since the last call, oldest first. */
RGAPI RgResult RGCONV rgExtServerPoll( RgExtServer server, RgExtFrameResult* pResults, uint32_t maxCount, uint32_t* pCount );
/* Blocks until no request is waiting or in flight. */
RGAPI RgResult RGCONV rgExtServerWaitIdle( RgExtServer server );
RGAPI RgResult RGCONV rgExtServerGetStats( RgExtServer server, RgExtServerStats* pOut );
/* The dispatch policy of the launches that follow (rgExtServerSetPolicy). */
typedef struct RgExtServerPolicy
{
    uint32_t schedule;                /* RG_EXT_SCHEDULE_*                                  */
    uint32_t batchFrames;             /* 1 .. the batchFrames the server was created with  */
    uint32_t framesInFlightPerStream; /* launches queued per stream, 1..8                  */
    uint32_t streams;                 /* streams that take new launches, 1 .. streamCount  */
    double   maxQueueSeconds;         /* 0 = requests do not expire                        */
} RgExtServerPolicy;

/* Changes the dispatch policy: fewer streams, frames in flight or frames per
 * launch bound the latency of served frames, more raise the throughput.  A
 * session with a frame in flight on a stream no longer in use finishes its
 * frames there before it moves. */
RGAPI RgResult RGCONV rgExtServerSetPolicy( RgExtServer server, const RgExtServerPolicy* pPolicy );
/* The session's framebuffer set: its own persistent planes (FINAL holds the
 * last finished frame) and views of its stream's within-frame planes, which
 * the stream's next frame overwrites.  Owned by the session. */
RGAPI RgResult RGCONV rgExtSessionGetFramebuffers( RgExtSession session, RgExtFramebuffers* pOut );

#ifdef __cplusplus
}
#endif

#endif /* RTGL1_EXT_H_ */
