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
enum { ORC_LIGHT_DISK = 0, ORC_LIGHT_POINT = 1, ORC_LIGHT_QUAD = 2, ORC_LIGHT_TRIANGLE = 3 };
enum { ORC_LS_UNIFORM = 0, ORC_LS_POWER = 1, ORC_LS_BVH = 2 };

typedef struct orc_material {
    uint32_t kind;        /* ORC_MAT_* */
    float base_color[3];
    float alpha[2];       /* GGX roughness; Dielectric alpha (1,1) = Lambert only */
} orc_material;

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
} orc_render_params;

typedef struct orc_scene orc_scene;

orc_scene* orc_scene_create(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                            const uint32_t* tri_mat, const orc_material* mats, uint32_t nmats,
                            const float* normals, const orc_light* lights, uint32_t nl,
                            int light_sampler, const orc_camera* cam, const float* bg3);
void orc_scene_destroy(orc_scene* s);
void orc_render(orc_scene* s, const orc_render_params* rp, uint32_t* states, double* accum,
                uint32_t nthreads);

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
