/* CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h): the ray-traced
 * passes of a frame in FP64 (oracle_pathtrace.cpp). */
#ifndef RG_ORACLE_PATHTRACE_H_
#define RG_ORACLE_PATHTRACE_H_

#include "oracle.h"
#include "oracle_denoise.h"
#include "oracle_texture.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One geometry (an uploaded primitive) as the frame sees it. */
typedef struct OrGeom
{
    float    prevFromCur[ 12 ]; /* row-major 3x4: previous model * model^-1 */
    uint32_t isStatic;          /* prevFromCur is the identity              */
    uint32_t color;             /* RgColor4DPacked32                        */
    float    emissive;
    float    roughness;
    float    metallic;
    uint32_t flags;             /* RgMeshPrimitiveFlags                     */
    uint32_t firstTri;
    uint32_t triCount;
    uint32_t texture;           /* 1 + index into OrPtScene::textures, 0: none */
    uint32_t texOrm, texNormal, texEmissive; /* the _orm, _n and _e maps, same ids */
} OrGeom;

/* World-space triangles with their geometries and the frame's lights. */
typedef struct OrPtScene
{
    const float*                        tris;     /* triCount * 9 */
    const float*                        normals;  /* triCount * 9 */
    const uint32_t*                     triGeom;  /* triCount     */
    uint32_t                            triCount;
    const OrGeom*                       geoms;
    const uint64_t*                     geomIds;  /* uniqueObjectID << 32 | primitive index */
    uint32_t                            geomCount;
    const RgDirectionalLightUploadInfo* dirLight; /* may be null */
    const RgSphericalLightUploadInfo*   spheres;
    uint32_t                            sphereCount;
    /* regular lights in this order: spheres, spots, polygons */
    const RgSpotLightUploadInfo*        spots;
    uint32_t                            spotCount;
    const RgPolygonalLightUploadInfo*   polys;
    uint32_t                            polyCount;
    /* Albedo textures and per-triangle texture coordinates (6 floats). */
    const struct OrTexture* const*      textures;
    uint32_t                            textureCount;
    const float*                        texcoords;
} OrPtScene;

/* Flat normals of one primitive (CmVertexPreprocess.comp with
 * VertexPreprocessPartial.inl:43-120): for each triangle in index order,
 * normalize(cross(p1 - p0, p2 - p0)) of its positions is written to its
 * three vertices, so a vertex shared by several triangles keeps the last
 * one's normal.  positions and normals hold vertexCount * 3 floats (normals
 * in and out: vertices no triangle uses keep theirs); indices null for a
 * non-indexed primitive. */
void or_flat_normals( const float* positions, uint32_t vertexCount, const uint32_t* indices, uint32_t indexCount, float* normals );

typedef struct OrPathTracer OrPathTracer;
OrPathTracer* or_pt_create( int width, int height );
void          or_pt_destroy( OrPathTracer* pt );
/* The traced passes of one frame into fb's current planes (previous-frame
 * planes are read from fb's prev planes); light grid and light / geometry
 * matches persist in pt.  p supplies the camera and settings; its light
 * counts are replaced by the scene's. */
void or_pt_frame( OrPathTracer* pt, const OrPtScene* scene, const OrBvh* bvh, const RgExtFrameParams* p, const OrFramebuffers* fb, int threads );

#ifdef __cplusplus
}
#endif

#endif
