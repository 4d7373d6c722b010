/* CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h): A-SVGF in FP64. */
#ifndef RG_ORACLE_DENOISE_H_
#define RG_ORACLE_DENOISE_H_

#include <stdint.h>

#include "RTGL1/RTGL1_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host planes of the framebuffer set, one double array per plane
 * (components interleaved); prev[] holds the previous-frame copy of planes
 * with a PREV flag (others may be null). */
typedef struct OrFramebuffers
{
    int     width;
    int     height;
    double* cur[ RG_EXT_PLANE_COUNT ];
    double* prev[ RG_EXT_PLANE_COUNT ];
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

#ifdef __cplusplus
}
#endif

#endif
