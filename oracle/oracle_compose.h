/* CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h): final compose in FP64. */
#ifndef RG_ORACLE_COMPOSE_H_
#define RG_ORACLE_COMPOSE_H_

#include "oracle.h"
#include "oracle_denoise.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Luma coefficients the LPM setup derives for Rec.709 / D65. */
void or_lpm_luma_coefficients( double out[ 3 ] );
/* One colour through the LPM 709->709 map. */
void or_lpm_map( const float saturation[ 3 ], const float crosstalk[ 3 ], const double in[ 3 ], double out[ 3 ] );
/* PRE_FINAL -> FINAL of fb (checkerboard resolve, auto exposure, LPM). */
/* FINAL = LPM(PRE_FINAL * exposure + SCREEN_EMISSION * emissionMaxScreenColor)
 * + ACID_FOG, the three resolved alike on split pixels. */
void or_compose_final( const OrFramebuffers* fb, const float saturation[ 3 ], const float crosstalk[ 3 ], double avgLuminance,
                       double emissionMaxScreenColor );

#ifdef __cplusplus
}
#endif

#endif
