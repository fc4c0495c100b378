// RGL-MI355X public C ABI.
//
// This header re-declares the C interface of RobotecGPULidar 0.19.0 so that engine plugins (Unity, O3DE,
// Gazebo) and anything else that binds `include/rgl/api/core.h` of the reference can load this library
// instead. Names, argument order, POD layouts, enum values and status codes are kept identical: they are
// the binary contract. Every declaration cites the reference declaration it replaces
// (file:line in RobotecGPULidar 0.19.0). The implementation behind it is HIP for gfx950 (MI355X).

#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <type_traits>
#define NO_MANGLING extern "C"
#else
#define NO_MANGLING
#endif

#if defined RGL_STATIC
#define RGL_VISIBLE
#else
#define RGL_VISIBLE __attribute__((visibility("default")))
#endif

#define RGL_API NO_MANGLING RGL_VISIBLE

// Version of the reference API this library is a drop-in for (include/rgl/api/core.h:52-54).
#define RGL_VERSION_MAJOR 0
#define RGL_VERSION_MINOR 19
#define RGL_VERSION_PATCH 0

// Entity id given to points that hit nothing (include/rgl/api/core.h:58).
#define RGL_ENTITY_INVALID_ID 0
// Entity id assigned when the user sets none: largest 28-bit integer (include/rgl/api/core.h:62).
#define RGL_DEFAULT_ENTITY_ID 268435455

// ---------------------------------------------------------------- POD types (core.h:67-119)

typedef struct
{
	float value[2];
} rgl_vec2f;

typedef struct
{
	float value[3];
} rgl_vec3f;

typedef struct
{
	int32_t value[3];
} rgl_vec3i;

// Row-major 3x4 affine transform; rays use column 3 as origin and column 2 as direction.
typedef struct
{
	float value[3][4];
} rgl_mat3x4f;

// Radar scope (include/rgl/api/core.h:124-150).
typedef struct
{
	float begin_distance;
	float end_distance;
	float distance_separation_threshold;
	float radial_speed_separation_threshold;
	float azimuth_separation_threshold;
} rgl_radar_scope_t;

// Bone weights for skinning (include/rgl/api/core.h:159-177).
typedef struct
{
	float weights[4];
	int32_t bone_indexes[4];
} rgl_bone_weights_t;

#ifdef __cplusplus
static_assert(sizeof(rgl_vec2f) == 2 * sizeof(float));
static_assert(sizeof(rgl_vec3f) == 3 * sizeof(float));
static_assert(sizeof(rgl_vec3i) == 3 * sizeof(int32_t));
static_assert(sizeof(rgl_mat3x4f) == 12 * sizeof(float));
static_assert(sizeof(rgl_radar_scope_t) == 5 * sizeof(float));
static_assert(sizeof(rgl_bone_weights_t) == 4 * sizeof(float) + 4 * sizeof(int32_t));
static_assert(std::is_trivial<rgl_mat3x4f>::value && std::is_standard_layout<rgl_mat3x4f>::value);
#endif

// Radar object class (include/rgl/api/core.h:182-193).
typedef enum : int32_t
{
	RGL_RADAR_CLASS_CAR = 0,
	RGL_RADAR_CLASS_TRUCK,
	RGL_RADAR_CLASS_MOTORCYCLE,
	RGL_RADAR_CLASS_BICYCLE,
	RGL_RADAR_CLASS_PEDESTRIAN,
	RGL_RADAR_CLASS_ANIMAL,
	RGL_RADAR_CLASS_HAZARD,
	RGL_RADAR_CLASS_UNKNOWN,
	RGL_RADAR_CLASS_COUNT
} rgl_radar_object_class_t;

// ---------------------------------------------------------------- opaque handles (core.h:199-224)
typedef struct Mesh* rgl_mesh_t;
typedef struct Entity* rgl_entity_t;
typedef struct Texture* rgl_texture_t;
typedef struct Node* rgl_node_t;
typedef struct Scene* rgl_scene_t;

// Extensions (include/rgl/api/core.h:231-238). None of them is built into this library.
typedef enum : int32_t
{
	RGL_EXTENSION_PCL = 0,
	RGL_EXTENSION_ROS2 = 1,
	RGL_EXTENSION_UDP = 2,
	RGL_EXTENSION_WEATHER = 3,
	RGL_EXTENSION_COUNT
} rgl_extension_t;

// Status codes (include/rgl/api/core.h:244-327). Unrecoverable: INVALID_STATE, LOGGING_ERROR,
// INITIALIZATION_ERROR, INTERNAL_EXCEPTION.
typedef enum : int32_t
{
	RGL_SUCCESS = 0,
	RGL_INVALID_ARGUMENT,
	RGL_INVALID_STATE,
	RGL_LOGGING_ERROR,
	RGL_INVALID_API_OBJECT,
	RGL_INVALID_FILE_PATH,
	RGL_TAPE_ERROR,
	RGL_UDP_ERROR,
	RGL_ROS2_ERROR,
	RGL_INVALID_PIPELINE,
	RGL_INITIALIZATION_ERROR,
	RGL_NOT_IMPLEMENTED = 404,
	RGL_INTERNAL_EXCEPTION = 500,
} rgl_status_t;

// Log levels (include/rgl/api/core.h:332-343).
typedef enum : int32_t
{
	RGL_LOG_LEVEL_ALL = 0,
	RGL_LOG_LEVEL_TRACE = 0,
	RGL_LOG_LEVEL_DEBUG = 1,
	RGL_LOG_LEVEL_INFO = 2,
	RGL_LOG_LEVEL_WARN = 3,
	RGL_LOG_LEVEL_ERROR = 4,
	RGL_LOG_LEVEL_CRITICAL = 5,
	RGL_LOG_LEVEL_OFF = 6,
	RGL_LOG_LEVEL_COUNT = 7
} rgl_log_level_t;

// Point fields (include/rgl/api/core.h:348-450). Values are the wire contract.
typedef enum : int32_t
{
	RGL_FIELD_XYZ_VEC3_F32 = 1,
	RGL_FIELD_INTENSITY_F32,
	RGL_FIELD_INTENSITY_U8,
	RGL_FIELD_IS_HIT_I32,
	RGL_FIELD_IS_GROUND_I32,
	RGL_FIELD_RAY_IDX_U32,
	RGL_FIELD_ENTITY_ID_I32,
	RGL_FIELD_DISTANCE_F32,
	RGL_FIELD_AZIMUTH_F32,
	RGL_FIELD_ELEVATION_F32,
	RGL_FIELD_RING_ID_U16,
	RGL_FIELD_RETURN_TYPE_U8,
	RGL_FIELD_TIME_STAMP_F64,
	RGL_FIELD_TIME_STAMP_U32,
	RGL_FIELD_ABSOLUTE_VELOCITY_VEC3_F32,
	RGL_FIELD_RELATIVE_VELOCITY_VEC3_F32,
	RGL_FIELD_RADIAL_SPEED_F32,
	RGL_FIELD_POWER_F32,
	RGL_FIELD_RCS_F32,
	RGL_FIELD_NOISE_F32,
	RGL_FIELD_SNR_F32,
	RGL_FIELD_NORMAL_VEC3_F32,
	RGL_FIELD_INCIDENT_ANGLE_F32,
	RGL_FIELD_RAY_POSE_MAT3x4_F32,
	RGL_FIELD_LASER_RETRO_F32,
	RGL_FIELD_PADDING_8 = 1024,
	RGL_FIELD_PADDING_16,
	RGL_FIELD_PADDING_32,
	RGL_FIELD_DYNAMIC_FORMAT = 13842,
} rgl_field_t;

// Return-mode bit codec (include/rgl/api/core.h:453-504): byte 3 = number of returns,
// byte r (r = 0, 1) = rgl_return_type_t of return r.
const int32_t RGL_RETURN_MODE_BIT_SHIFT = 24;
const int32_t RGL_RETURN_MODE_SINGLE = 1 << RGL_RETURN_MODE_BIT_SHIFT;
const int32_t RGL_RETURN_MODE_DUAL = 2 << RGL_RETURN_MODE_BIT_SHIFT;
const int32_t RGL_RETURN_TYPE_BIT_SHIFT = 8;

typedef enum : uint8_t
{
	RGL_RETURN_TYPE_UNKNOWN = 0,
	RGL_RETURN_TYPE_STRONGEST = 1,
	RGL_RETURN_TYPE_LAST = 2,
	RGL_RETURN_TYPE_SECOND = 3,
	RGL_RETURN_TYPE_FIRST = 4,
	RGL_RETURN_TYPE_SECOND_STRONGEST = 5,
} rgl_return_type_t;

typedef enum : int32_t
{
	RGL_RETURN_UNKNOWN = RGL_RETURN_TYPE_UNKNOWN,
	RGL_RETURN_FIRST = RGL_RETURN_MODE_SINGLE | RGL_RETURN_TYPE_FIRST,
	RGL_RETURN_SECOND = RGL_RETURN_MODE_SINGLE | RGL_RETURN_TYPE_SECOND,
	RGL_RETURN_LAST = RGL_RETURN_MODE_SINGLE | RGL_RETURN_TYPE_LAST,
	RGL_RETURN_STRONGEST = RGL_RETURN_MODE_SINGLE | RGL_RETURN_TYPE_STRONGEST,
	RGL_RETURN_LAST_STRONGEST = RGL_RETURN_MODE_DUAL | RGL_RETURN_TYPE_LAST |
	                            (RGL_RETURN_TYPE_STRONGEST << RGL_RETURN_TYPE_BIT_SHIFT),
	RGL_RETURN_FIRST_LAST = RGL_RETURN_MODE_DUAL | RGL_RETURN_TYPE_FIRST |
	                        (RGL_RETURN_TYPE_LAST << RGL_RETURN_TYPE_BIT_SHIFT),
	RGL_RETURN_FIRST_STRONGEST = RGL_RETURN_MODE_DUAL | RGL_RETURN_TYPE_FIRST |
	                             (RGL_RETURN_TYPE_STRONGEST << RGL_RETURN_TYPE_BIT_SHIFT),
	RGL_RETURN_STRONGEST_SECOND_STRONGEST = RGL_RETURN_MODE_DUAL | RGL_RETURN_TYPE_STRONGEST |
	                                        (RGL_RETURN_TYPE_SECOND_STRONGEST << RGL_RETURN_TYPE_BIT_SHIFT),
	RGL_RETURN_FIRST_SECOND = RGL_RETURN_MODE_DUAL | RGL_RETURN_TYPE_FIRST |
	                          (RGL_RETURN_TYPE_SECOND << RGL_RETURN_TYPE_BIT_SHIFT),
} rgl_return_mode_t;

// Rotation axis selector (include/rgl/api/core.h:509-514).
typedef enum : int32_t
{
	RGL_AXIS_X = 1,
	RGL_AXIS_Y = 2,
	RGL_AXIS_Z = 3,
} rgl_axis_t;

// ---------------------------------------------------------------- general (core.h:526-560)
RGL_API rgl_status_t rgl_get_version_info(int32_t* out_major, int32_t* out_minor, int32_t* out_patch); // core.h:526
RGL_API rgl_status_t rgl_get_extension_info(rgl_extension_t extension, int32_t* out_available);      // core.h:534
RGL_API rgl_status_t rgl_configure_logging(rgl_log_level_t log_level, const char* log_file_path,
                                           bool use_stdout);                                    // core.h:546
RGL_API void rgl_get_last_error_string(const char** out_error_string);                          // core.h:553
RGL_API rgl_status_t rgl_cleanup(void);                                                         // core.h:560

// ---------------------------------------------------------------- mesh (core.h:573-620)
RGL_API rgl_status_t rgl_mesh_create(rgl_mesh_t* out_mesh, const rgl_vec3f* vertices, int32_t vertex_count,
                                     const rgl_vec3i* indices, int32_t index_count);            // core.h:573
RGL_API rgl_status_t rgl_mesh_set_texture_coords(rgl_mesh_t mesh, const rgl_vec2f* uvs, int32_t uv_count); // core.h:583
RGL_API rgl_status_t rgl_mesh_set_bone_weights(rgl_mesh_t mesh, const rgl_bone_weights_t* bone_weights,
                                               int32_t bone_weights_count);                    // core.h:593
RGL_API rgl_status_t rgl_mesh_set_restposes(rgl_mesh_t mesh, const rgl_mat3x4f* restposes,
                                            int32_t bones_count);                              // core.h:605
RGL_API rgl_status_t rgl_mesh_destroy(rgl_mesh_t mesh);                                         // core.h:612
RGL_API rgl_status_t rgl_mesh_is_alive(rgl_mesh_t mesh, bool* out_alive);                       // core.h:620

// ---------------------------------------------------------------- entity (core.h:631-701)
RGL_API rgl_status_t rgl_entity_create(rgl_entity_t* out_entity, rgl_scene_t scene, rgl_mesh_t mesh); // core.h:631
RGL_API rgl_status_t rgl_entity_destroy(rgl_entity_t entity);                                         // core.h:638
RGL_API rgl_status_t rgl_entity_set_transform(rgl_entity_t entity, const rgl_mat3x4f* transform);     // core.h:646
RGL_API rgl_status_t rgl_entity_set_pose_world(rgl_entity_t entity, const rgl_mat3x4f* pose,
                                               int32_t bones_count);                   // core.h:659
RGL_API rgl_status_t rgl_entity_set_id(rgl_entity_t entity, int32_t id);                              // core.h:666
RGL_API rgl_status_t rgl_entity_set_intensity_texture(rgl_entity_t entity, rgl_texture_t texture);    // core.h:673
RGL_API rgl_status_t rgl_entity_set_laser_retro(rgl_entity_t entity, float retro);                    // core.h:682
RGL_API rgl_status_t rgl_entity_apply_external_animation(rgl_entity_t entity, const rgl_vec3f* vertices,
                                                         int32_t vertex_count);                       // core.h:693
RGL_API rgl_status_t rgl_entity_is_alive(rgl_entity_t entity, bool* out_alive);                       // core.h:701

// ---------------------------------------------------------------- texture (core.h:713-728)
RGL_API rgl_status_t rgl_texture_create(rgl_texture_t* out_texture, const void* texels, int32_t width,
                                        int32_t height);                                      // core.h:713
RGL_API rgl_status_t rgl_texture_destroy(rgl_texture_t texture);                               // core.h:720
RGL_API rgl_status_t rgl_texture_is_alive(rgl_texture_t texture, bool* out_alive);             // core.h:728

// ---------------------------------------------------------------- scene (core.h:740)
RGL_API rgl_status_t rgl_scene_set_time(rgl_scene_t scene, uint64_t nanoseconds); // core.h:740

// ---------------------------------------------------------------- rays nodes (core.h:754-801)
RGL_API rgl_status_t rgl_node_rays_from_mat3x4f(rgl_node_t* node, const rgl_mat3x4f* rays, int32_t ray_count); // core.h:754
RGL_API rgl_status_t rgl_node_rays_set_ring_ids(rgl_node_t* node, const int32_t* ring_ids,
                                                int32_t ring_ids_count);                                       // core.h:765
RGL_API rgl_status_t rgl_node_rays_set_range(rgl_node_t* node, const rgl_vec2f* ranges, int32_t ranges_count); // core.h:779
RGL_API rgl_status_t rgl_node_rays_set_time_offsets(rgl_node_t* node, const float* offsets,
                                                    int32_t offsets_count);                                    // core.h:790
RGL_API rgl_status_t rgl_node_rays_transform(rgl_node_t* node, const rgl_mat3x4f* transform);                  // core.h:801

// ---------------------------------------------------------------- points / raytrace nodes (core.h:813-940)
RGL_API rgl_status_t rgl_node_points_transform(rgl_node_t* node, const rgl_mat3x4f* transform); // core.h:813
RGL_API rgl_status_t rgl_node_raytrace(rgl_node_t* node, rgl_scene_t scene);                    // core.h:825
RGL_API rgl_status_t rgl_node_raytrace_configure_velocity(rgl_node_t node, const rgl_vec3f* linear_velocity,
                                                          const rgl_vec3f* angular_velocity);  // core.h:837
RGL_API rgl_status_t rgl_node_raytrace_configure_distortion(rgl_node_t node, bool enable);     // core.h:849
RGL_API rgl_status_t rgl_node_raytrace_configure_non_hits(rgl_node_t node, float nearDistance,
                                                          float farDistance);                  // core.h:862
RGL_API rgl_status_t rgl_node_raytrace_configure_mask(rgl_node_t node, const int8_t* rays_mask,
                                                      int32_t rays_count);                     // core.h:871
RGL_API rgl_status_t rgl_node_raytrace_configure_beam_divergence(rgl_node_t node, float horizontal_beam_divergence,
                                                                 float vertical_beam_divergence); // core.h:882
RGL_API rgl_status_t rgl_node_raytrace_configure_default_intensity(rgl_node_t node, float default_intensity); // core.h:893
RGL_API rgl_status_t rgl_node_raytrace_configure_return_mode(rgl_node_t node, rgl_return_mode_t return_mode); // core.h:905
RGL_API rgl_status_t rgl_node_points_format(rgl_node_t* node, const rgl_field_t* fields, int32_t field_count); // core.h:917
RGL_API rgl_status_t rgl_node_points_yield(rgl_node_t* node, const rgl_field_t* fields, int32_t field_count);  // core.h:928
RGL_API rgl_status_t rgl_node_points_compact_by_field(rgl_node_t* node, rgl_field_t field);                   // core.h:940
// Alias of rgl_node_points_compact_by_field(node, RGL_FIELD_IS_HIT_I32); the reference removed this name in 0.18.0
// (CHANGELOG.md:80-83). Kept for callers that still bind it.
RGL_API rgl_status_t rgl_node_points_compact(rgl_node_t* node);

RGL_API rgl_status_t rgl_node_points_spatial_merge(rgl_node_t* node, const rgl_field_t* fields,
                                                   int32_t field_count);          // core.h:957 (not implemented)
RGL_API rgl_status_t rgl_node_points_temporal_merge(rgl_node_t* node, const rgl_field_t* fields,
                                                    int32_t field_count);         // core.h:971 (not implemented)
RGL_API rgl_status_t rgl_node_points_from_array(rgl_node_t* node, const void* points, int32_t points_count,
                                                const rgl_field_t* fields, int32_t field_count); // core.h:995
RGL_API rgl_status_t rgl_node_points_radar_postprocess(rgl_node_t* node, const rgl_radar_scope_t* radar_scopes,
                                                       int32_t radar_scopes_count, float ray_azimuth_step,
                                                       float ray_elevation_step, float frequency, float power_transmitted,
                                                       float cumulative_device_gain, float received_noise_mean,
                                                       float received_noise_st_dev); // core.h:1020
RGL_API rgl_status_t rgl_node_points_radar_track_objects(rgl_node_t* node, float object_distance_threshold,
                                                         float object_azimuth_threshold, float object_elevation_threshold,
                                                         float object_radial_speed_threshold, float max_matching_distance,
                                                         float max_prediction_time_frame,
                                                         float movement_sensitivity); // core.h:1045 (not implemented)
RGL_API rgl_status_t rgl_node_points_radar_set_classes(rgl_node_t node, const int32_t* entity_ids,
                                                       const rgl_radar_object_class_t* object_classes,
                                                       int32_t count); // core.h:1060 (not implemented)
RGL_API rgl_status_t rgl_node_points_filter_ground(rgl_node_t* node, const rgl_vec3f* sensor_up_vector,
                                                   float ground_angle_threshold); // core.h:1073 (not implemented)

// ---------------------------------------------------------------- noise nodes (core.h:1087-1121)
RGL_API rgl_status_t rgl_node_gaussian_noise_angular_ray(rgl_node_t* node, float mean, float st_dev,
                                                         rgl_axis_t rotation_axis); // core.h:1087
RGL_API rgl_status_t rgl_node_gaussian_noise_angular_hitpoint(rgl_node_t* node, float mean, float st_dev,
                                                              rgl_axis_t rotation_axis); // core.h:1103
RGL_API rgl_status_t rgl_node_gaussian_noise_distance(rgl_node_t* node, float mean, float st_dev_base,
                                                      float st_dev_rise_per_meter); // core.h:1120

// ---------------------------------------------------------------- graph (core.h:1129-1198)
RGL_API rgl_status_t rgl_node_is_alive(rgl_node_t node, bool* out_alive);                          // core.h:1129
RGL_API rgl_status_t rgl_graph_run(rgl_node_t node);                                               // core.h:1138
RGL_API rgl_status_t rgl_graph_destroy(rgl_node_t node);                                           // core.h:1144
RGL_API rgl_status_t rgl_graph_get_result_size(rgl_node_t node, rgl_field_t field, int32_t* out_count,
                                               int32_t* out_size_of);                             // core.h:1155
RGL_API rgl_status_t rgl_graph_get_result_data(rgl_node_t node, rgl_field_t field, void* data);    // core.h:1164
RGL_API rgl_status_t rgl_graph_node_add_child(rgl_node_t parent, rgl_node_t child);                // core.h:1171
RGL_API rgl_status_t rgl_graph_node_remove_child(rgl_node_t parent, rgl_node_t child);             // core.h:1178
RGL_API rgl_status_t rgl_graph_node_set_priority(rgl_node_t node, int32_t priority);               // core.h:1191
RGL_API rgl_status_t rgl_graph_node_get_priority(rgl_node_t node, int32_t* out_priority);          // core.h:1198
