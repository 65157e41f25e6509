/*
 * thermal_api_compat.h -- the reference's external C ABI, served by libirtt.so.
 *
 * Same symbol names, argument meaning and struct layout as
 * src/implementations/thermal_renderer/thermal_api.hpp:10-62 and
 * thermal_api.cpp:25-45,300-536 of the reference, so a host that binds the
 * reference's thermal_renderer_lib (its own comment names a CAD "rhino
 * interface", thermal_api.hpp:18) can load this library instead.  One scene at
 * a time, process-global state, not thread-safe, load() first -- as in the
 * reference.
 *
 * Differences (each a fix, none changes a successful call's meaning):
 *  - every function returns 0 on success and a negative IRTT_ERR_* code on
 *    failure (the reference always returns 0); irtt_last_error() has the text;
 *  - simulate() builds the transport operator before the first step and after
 *    geometry changes (the reference's library path never does,
 *    thermal_renderer.cpp:391 is GUI-only, so it solves with T = 0);
 *  - load_sky() accepts any number of quads (the reference assumes the 145
 *    Tregenza patches, thermal_data.hpp:27,41,59-60).
 */
#ifndef THERMAL_API_COMPAT_H
#define THERMAL_API_COMPAT_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* thermal_api.cpp:25-45; sizeof == 52 on x86-64 (bools padded as by the
 * reference's compiler) */
typedef struct ObjectProperties {
    unsigned int vertex_offset;   /* in vertices (x3 floats) */
    unsigned int vertex_count;
    unsigned int indices_offset;  /* in triangles (x3 indices) */
    unsigned int indices_count;
    float kelvin;
    bool temperature_fixed;
    float thickness;
    float density;
    float heat_capacity;
    float heat_conductivity;
    float diffuse_reflectance;
    float specular_reflectance;
    bool diffuse_emission;
    bool traceable;
} ObjectProperties;

/* thermal_api.cpp:300-332 / 340-345 */
int load(bool _withConsole);
int unload(void);

/* thermal_api.cpp:334-358 */
int test_load_scene(void);
int test_lib(int _value);                       /* returns 2 * _value */
int test_fill_array(float *_arr, int _size);    /* _arr[i] *= 3 i */

/* thermal_api.cpp:360-406: per-object vertex and triangle ranges; triangle
 * indices are local to their object */
int load_geometry(float *_vertices, float *_vertex_colors, unsigned int _total_vertex_count,
                  unsigned int *_indices, unsigned int _total_indices_count,
                  ObjectProperties *_object_properties, unsigned int _object_count);

/* thermal_api.cpp:416-447 and 166-296: sky quads (4 indices each), values in
 * W/m^2; quads are split into (c,b,d),(b,a,d) and re-projected around the
 * scene bounding sphere; loading the sky loads the scene */
int load_sky(float *_vertices, unsigned int _vertex_count, unsigned int *_indices, float *_values,
             unsigned int _quad_count);

int unload_scene(void);                                          /* :408-414 */
int update_sky_values(float *_values, unsigned int _quad_count); /* :449-458 */
int reset_simulation(void);                                      /* :485-489 */
/* thermal_renderer.hpp:88: threads per triangle, rays per thread */
int set_ray_batch_count(unsigned int _ray_count, unsigned int _batch_count);
int set_minimum_sky_kelvin(float _min);                          /* :498-503 */
int set_steady_state(bool _enabled);                             /* :505-509 */
/* thermal_api.cpp:460-483: _time_step_count steps of _step_size_hours;
 * *_time_hours in: start time, out: end time */
int simulate(float _step_size_hours, unsigned int _time_step_count, float *_time_hours);
/* thermal_renderer.cpp:402-427: one value per vertex, written three times
 * (_total_vertex_count counts floats); _type 0 kelvin, 1 sigma*A*T^4 (W) */
int get_vertex_temperatures(float *_vertex_temperatures, unsigned int _total_vertex_count);
int get_vertex_values(float *_vertex_temperatures, unsigned int _total_vertex_count,
                      unsigned int _type);

#ifdef __cplusplus
}
#endif
#endif
