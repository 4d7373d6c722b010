/* CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h): A-SVGF in FP64. */
#ifndef RG_ORACLE_DENOISE_H_
#define RG_ORACLE_DENOISE_H_

#include <stdint.h>

#include "RTGL1/RTGL1_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

/* An oracle-only plane after the product's planes (C f64, PREV): 1 where
 * the pixel's ReSTIR GI reservoir depends on a decision that FP32 and FP64
 * arithmetic can take differently, else 0.  Such a decision is the sign of
 * the cosine at a spatial neighbour's sample point towards this pixel's
 * surface when the surface lies in the sample's own plane (flat-shaded
 * geometry: the pixel and the neighbour's bounce hit on one triangle's
 * plane), which decides whether the sample counts (RtRaygenIndirect.inl:
 * 365-403); a pixel whose temporal sample comes from such a pixel of the
 * previous frame inherits the mark. */
#define OR_PLANE_GI_TIE RG_EXT_PLANE_COUNT
#define OR_PLANE_COUNT ( RG_EXT_PLANE_COUNT + 1 )

/* Host planes of the framebuffer set, one double array per plane
 * (components interleaved); prev[] holds the previous-frame copy of planes
 * with a PREV flag (others may be null). */
typedef struct OrFramebuffers
{
    int     width;
    int     height;
    double* cur[ OR_PLANE_COUNT ];
    double* prev[ OR_PLANE_COUNT ];
} OrFramebuffers;

int or_plane_components( int plane );
/* 0 = checkerboard W x H, 1 = regular W x H, 2 = gradient strata
 * max(1,(W+1)/3) x max(1,(H+1)/3) (Source/Framebuffers.cpp:665-699). */
int or_plane_layout( int plane );
int or_plane_has_prev( int plane );

void or_denoise_gradient_atrous( const OrFramebuffers* fb, const RgExtFrameParams* p, uint32_t iteration );
void or_denoise_temporal( const OrFramebuffers* fb, const RgExtFrameParams* p );
void or_denoise_antifirefly( const OrFramebuffers* fb, const RgExtFrameParams* p );
void or_denoise_variance( const OrFramebuffers* fb, const RgExtFrameParams* p );
void or_denoise_atrous( const OrFramebuffers* fb, const RgExtFrameParams* p, uint32_t iteration );
void or_denoise( const OrFramebuffers* fb, const RgExtFrameParams* p );
/* Host threads the passes use (rows in parallel; the results do not depend
 * on the count).  1 by default. */
void or_denoise_set_threads( int threads );

#ifdef __cplusplus
}
#endif

#endif
