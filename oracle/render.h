/*
 * ORACLE -- test infrastructure only.  Render-path restatement (render.c).
 */
#ifndef ALPINE_ORACLE_RENDER_H
#define ALPINE_ORACLE_RENDER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MAT_DIELECTRIC = 0, ORC_MAT_CONDUCTOR = 1 };
enum { ORC_LIGHT_DISK = 0, ORC_LIGHT_POINT = 1, ORC_LIGHT_QUAD = 2, ORC_LIGHT_TRIANGLE = 3,
       ORC_LIGHT_POINT_INTENSITY = 4 };
enum { ORC_LS_UNIFORM = 0, ORC_LS_POWER = 1, ORC_LS_BVH = 2 };

typedef struct orc_material {
    uint32_t kind;        /* ORC_MAT_* */
    float base_color[3];
    float alpha[2];       /* GGX roughness; Dielectric alpha (1,1) = Lambert only */
    uint32_t base_color_tex;  /* 0: base_color; k: texture k - 1 */
    uint32_t normal_tex;      /* 0: none; k: texture k - 1 */
} orc_material;

/* Shading attributes expanded per triangle (the three corners' values, in corner order). */
typedef struct orc_shading {
    const float* tri_normals;    /* nt x 3 x 3, or NULL: ns = ng */
    const float* tri_tangents;   /* nt x 3 x 4 (xyz, handedness), or NULL */
    const float* tri_uvs;        /* nt x 3 x 2, or NULL: uv = (0, 0) */
    uint32_t ntargets;           /* morph targets */
    const float* tri_dnormals;   /* ntargets x nt x 3 x 3 normal deltas */
    const float* tri_dtangents;  /* ntargets x nt x 3 x 3 tangent deltas */
    const float* weights;        /* ntargets */
} orc_shading;

typedef struct orc_light {
    uint32_t kind;        /* ORC_LIGHT_* */
    float p0[3];          /* disk/point: position; quad/triangle: first corner */
    float p1[3], p2[3];   /* quad: p0 + e0, p0 + e1; triangle: the other two vertices */
    float normal[3];      /* disk */
    float color[3];       /* disk/point: colour; quad/triangle: emitted radiance */
    float power;          /* disk/point */
    float radius;         /* disk */
} orc_light;

typedef struct orc_camera {
    float eye[3], target[3], up[3];
    float fovy, aspect;
} orc_camera;

typedef struct orc_render_params {
    uint32_t width, height, spp, max_depth;
    uint32_t rr_depth;    /* 0 = no Russian roulette */
    /* Checks of the integrator's own tests; 0 for the reference's behaviour.
     * mis_mode: ORC_MIS_LIGHT_ONLY / ORC_MIS_BSDF_ONLY keep one of the two direct-light
     * strategies of estimateDirectIllumination (alpine.cpp:396-428) with weight 1.
     * plant: ORC_PLANT_* mistakes, so tests can show their closed forms catch them. */
    uint32_t mis_mode;
    uint32_t plant;
    /* crop_w, crop_h > 0: only the pixels [crop_x, crop_x + crop_w) x [crop_y, crop_y + crop_h)
     * of the width x height frame are traced; states and accum then hold those pixels only,
     * row-major within the crop (each still seeded by the caller with its frame index). */
    uint32_t crop_x, crop_y, crop_w, crop_h;
} orc_render_params;
#define ORC_MIS_BOTH 0u
#define ORC_MIS_LIGHT_ONLY 1u
#define ORC_MIS_BSDF_ONLY 2u
#define ORC_PLANT_MIS_WEIGHT 1u     /* BSDF strategy weighted 1 instead of the power heuristic */
#define ORC_PLANT_NO_NG_KILL 2u     /* continuation not killed when dot(wi, ng) <= 0 */
#define ORC_PLANT_SHADOW_NG 4u      /* shadow rays offset along ng instead of ns */
#define ORC_PLANT_NO_THROUGHPUT 8u  /* path throughput not multiplied by the BSDF estimator */

typedef struct orc_scene orc_scene;

orc_scene* orc_scene_create(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                            const uint32_t* tri_mat, const orc_material* mats, uint32_t nmats,
                            const float* normals, const orc_light* lights, uint32_t nl,
                            int light_sampler, const orc_camera* cam, const float* bg3);
void orc_scene_destroy(orc_scene* s);
/* aov (optional): width x height x 6 doubles, first-hit base colour and shading normal sums. */
void orc_render(orc_scene* s, const orc_render_params* rp, uint32_t* states, double* accum,
                double* aov, uint32_t nthreads);
/* Replace the scene's shading attributes (overrides orc_scene_create's normals). */
void orc_scene_set_shading(orc_scene* s, const orc_shading* sh);
/* texels: all textures' float4 texels back to back; sizes: width, height per texture. */
void orc_scene_set_textures(orc_scene* s, const float* texels, const uint32_t* sizes, uint32_t ntex);
void orc_scene_set_environment(orc_scene* s, const float* rgb, uint32_t width, uint32_t height);
/* Mesh::getIntersectionAttributes of triangle prim at barycentrics (b1, b2): ns, ss, ts, uv,
 * and the material's base colour there. */
void orc_shading_attributes(const orc_scene* s, uint32_t prim, float b1, float b2, float* ns3,
                            float* ss3, float* ts3, float* uv2, float* base_color3);
void orc_environment_sample(const orc_scene* s, const float* dir3, float* rgb3);

void orc_sampler_reset(uint32_t* state4, uint32_t seed);
void orc_postprocess(const float* accum, uint32_t n_pixels, uint32_t total_spp, int tonemap,
                     uint8_t* out);
float orc_sampler_get1d(uint32_t* state4);
void orc_camera_ray(const orc_scene* s, float x, float y, float* o3, float* d3);
int orc_mat_sample(const orc_material* m, const float* wo3, const float* u2, float* est3,
                   float* wi3, float* pdf);
void orc_mat_eval(const orc_material* m, const float* wo3, const float* wi3, float* f3,
                  float* pdf);
int orc_light_sampler_sample(const orc_scene* s, float u, const float* hit, const float* ns,
                             float* pdf);
float orc_light_sampler_pdf(const orc_scene* s, int light, const float* hit, const float* ns);
int orc_light_sample(const orc_scene* s, int light, const float* u2, const float* hit,
                     float* le3, float* wi3, float* dist, float* pdf);
float orc_light_pdf(const orc_scene* s, int light, const float* hit, const float* wi3,
                    float* dist);

#ifdef __cplusplus
}
#endif

#endif
