/* CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h): albedo textures in
 * FP64 (oracle_texture.cpp).  Textures of rgProvideOriginalTexture
 * (Source/TextureManager.cpp:398-440) are R8G8B8A8 sRGB with a mip chain
 * made by linear blits (TextureUploader.cpp:104-175); sampling follows the
 * Vulkan rules for the reference's samplers (SamplerManager.cpp:213-265). */
#ifndef RG_ORACLE_TEXTURE_H_
#define RG_ORACLE_TEXTURE_H_

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrTexture OrTexture;

/* filter / address modes: RgSamplerFilter, RgSamplerAddressMode values. */
/* srgb: 0 for the UNORM maps (occlusion-roughness-metallic, normal). */
OrTexture* or_texture_create( const uint8_t* rgba, uint32_t width, uint32_t height, int filter, int addressU, int addressV,
                              int forceMinificationLinear, int srgb );
void       or_texture_free( OrTexture* t );
uint32_t   or_texture_mip_count( const OrTexture* t );
/* Level l's texels (RGBA8) into out (width_l * height_l * 4 bytes). */
void       or_texture_copy_mip( const OrTexture* t, uint32_t level, uint8_t* out );
/* textureLod and textureGrad: linear RGBA (sRGB-decoded colour). */
void       or_texture_sample_lod( const OrTexture* t, double u, double v, double lod, double out[ 4 ] );
void       or_texture_sample_grad( const OrTexture* t, double u, double v, double dudx, double dvdx, double dudy, double dvdy,
                                   double out[ 4 ] );

#ifdef __cplusplus
}
#endif

#endif
