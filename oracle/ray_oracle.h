/*
 * ray_oracle.h — TEST INFRASTRUCTURE: the CPU oracle.
 *
 * A scalar C++ restatement of the Ray library's CPU "Ref" backend for the
 * wavefront path (generate -> trace -> sort -> shade/NEE -> accumulate), used
 * only by the tests, the smoke check and bench.py's CPU-baseline leg as the
 * checker of the HIP kernels.  It is pinned against the compiled reference
 * (oracle/_ref/libray_ref.so, built from the reference's own sources) and the
 * reference's golden vectors; see tests/test_oracle.py.
 *
 * Scene data comes in as an rh_scene_view (include/rh_types.h).
 */
#ifndef RAY_ORACLE_H
#define RAY_ORACLE_H

#include "rh_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* visit counters of the traversal: node visits, leaf triangle tests, instance tests */
typedef struct orc_counters {
    unsigned long long nodes, tris, instances;
} orc_counters;

int orc_halton_perms(uint16_t *out, size_t count);
int orc_halton_table(int iteration, float *out);

int orc_gen_primary(const rh_camera *cam, const int32_t rect[4], int w, int h, int iteration, const float *halton,
                    rh_ray_packet *out);

int orc_trace(const rh_scene_view *sc, const rh_ray_packet *rays, size_t n, int any_hit, const float *tmax,
              rh_hit *out, orc_counters *counters);

int orc_ray_hashes(const rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3],
                   uint32_t *out);
int orc_sort_rays(rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3]);

int orc_shade(const rh_scene_view *sc, const rh_pass_info *pi, const rh_ray_packet *rays, const rh_hit *hits,
              size_t n, const float *halton, float *out_rgba, rh_ray_packet *out_secondary, int *out_count,
              int *out_src);

/* Runs `passes` RenderScene passes over `rect` of a w x h image, continuing
 * from region iteration *iteration (updated), accumulating into clean (w*h*4
 * floats, caller-owned, zero to start) and writing the tonemapped result to
 * final_rgba.  ray_counts: primary, secondary, shadow. */
/* FP64 gradient of sum_pixels adj . (mean image of passes 1..passes) with
 * respect to every material's main_color (3 doubles per material, added to
 * grad) and, when texel_grad is not null, every stored texel's value in
 * [0, 1] (4 doubles per texel of the scene's texel array, alpha unused), by
 * path replay of the same passes; scale multiplies each pass's term (1/passes
 * for the mean). */
int orc_adjoint(const rh_scene_view *sc, const rh_camera *cam, int w, int h, const int32_t rect[4], int passes,
                const float *adj_rgba, double scale, double *grad, double *texel_grad);
/* the max_total_depth clamp (RH_MAX_BOUNCES by default, up to RH_MAX_BOUNCES_EXT) */
int orc_set_max_bounces(int n);
int orc_render(const rh_scene_view *sc, const rh_camera *cam, int w, int h, const int32_t rect[4], int passes,
               int *iteration, float *clean, float *final_rgba, unsigned long long ray_counts[3]);
/* orc_render with the OutputSH data: sh_temp and sh_clean hold 12 floats per
 * pixel (zeroed by the caller before the first pass, kept between calls) */
int orc_render_sh(const rh_scene_view *sc, const rh_camera *cam, int w, int h, const int32_t rect[4], int passes,
                  int *iteration, float *clean, float *final_rgba, unsigned long long ray_counts[3], float *sh_temp,
                  float *sh_clean);

#ifdef __cplusplus
}
#endif

#endif
