// RGL-MI355X extensions: entry points with no counterpart in the reference API.
//
// They expose what the MI355X implementation adds on top of RGL's C interface: control of the seeded noise
// stream (the reference seeds noise from std::random_device, src/graph/NodesCore.hpp:536,563,590, so it cannot
// be replayed), the TLAS update policy for moving entities, and introspection used by tests and benchmarks.
#pragma once

#include <rgl/api/core.h>

// Sets the seed of a Gaussian noise node's stream and restarts its run counter. The noise drawn for element i in
// run r is a pure function of (seed, r, node kind, i), so the same seed replays the same noise on CPU and GPU.
RGL_API rgl_status_t rgl_node_gaussian_noise_set_seed(rgl_node_t node, uint64_t seed);

// Reports the stream identity (seed, next run index, node-kind tag) of a Gaussian noise node.
RGL_API rgl_status_t rgl_node_gaussian_noise_get_stream(rgl_node_t node, uint64_t* out_seed, uint32_t* out_run,
                                                        uint32_t* out_tag);

// Chooses how the scene's TLAS follows entity transform changes when no entity was added or removed:
// refit = true keeps the topology and refits boxes; false rebuilds the TLAS (default).
RGL_API rgl_status_t rgl_scene_set_tlas_refit(rgl_scene_t scene, bool refit);

// BVH node format used by traversal in every scene: 0 = binary with FP32 child boxes (64 B nodes), 1 = 4-wide with
// FP32 child boxes (128 B), 2 = 4-wide with child boxes quantized to 8 bits relative to the node box and rounded
// outward (64 B). All formats index the same leaves, so results are identical; they differ in speed only. Scenes are
// rebuilt on their next use.
RGL_API rgl_status_t rgl_configure_bvh_format(int32_t format);

// Reports the BVH node format in use.
RGL_API rgl_status_t rgl_get_bvh_format(int32_t* out_format);

// Per-ray traversal counters of a raytrace node's last configuration: internal BVH nodes visited, triangles tested
// and instances entered. Arrays hold `ray_count` entries (the node's ray count).
RGL_API rgl_status_t rgl_node_raytrace_traversal_stats(rgl_node_t node, int32_t* nodes_visited, int32_t* tris_tested,
                                                       int32_t* instances_entered, int32_t ray_count);

// Closest hit of every primary ray by testing all triangles of all instances (no BVH). Misses give t = -1,
// instance = -1, triangle = -1. Intended for verifying the BVH traversal.
RGL_API rgl_status_t rgl_node_raytrace_brute_force(rgl_node_t node, float* out_t, int32_t* out_instance,
                                                   int32_t* out_prim, int32_t ray_count);

// Enables (and resets) HIP-event timing of a raytrace node's kernel launches on its graph stream.
RGL_API rgl_status_t rgl_node_raytrace_set_kernel_timing(rgl_node_t node, bool enable);

// Total device time (ms) and number of raytrace kernel launches timed since timing was enabled; waits for them.
RGL_API rgl_status_t rgl_node_raytrace_get_kernel_timing(rgl_node_t node, double* out_total_ms, int32_t* out_launches);

// Creates an empty scene. Entities created with it raytrace only against it; rgl_scene_t NULL stays the default
// scene of the reference API (rgl_entity_create, include/rgl/api/core.h:631, and rgl_node_raytrace take a scene
// argument, but the reference accepts only NULL there: src/api/apiCore.cpp:297,922).
RGL_API rgl_status_t rgl_scene_create(rgl_scene_t* out_scene);

// Destroys a scene created with rgl_scene_create together with its entities; waits for graphs in flight.
RGL_API rgl_status_t rgl_scene_destroy(rgl_scene_t scene);

// Runs the graphs containing `nodes[0..node_count)` (any node of each graph; a graph named twice runs once) as one
// batch: nodes of the same kind at the same depth of every graph execute in one launch. Each graph's results are
// identical to those of rgl_graph_run on the same graph and state; results are read with
// rgl_graph_get_result_size / rgl_graph_get_result_data as usual. Graphs may raytrace different scenes.
RGL_API rgl_status_t rgl_graph_run_batch(const rgl_node_t* nodes, int32_t node_count);

// Enables (and resets) HIP-event timing of the batched raytrace launches.
RGL_API rgl_status_t rgl_graph_batch_set_kernel_timing(bool enable);

// Total device time (ms) and number of batched raytrace launches timed since timing was enabled; waits for them.
RGL_API rgl_status_t rgl_graph_batch_get_kernel_timing(double* out_total_ms, int32_t* out_launches);

// Standard normal variates normal(seed, run, tag, i) for i in [0, count), computed on the GPU.
RGL_API rgl_status_t rgl_noise_normals(uint64_t seed, uint32_t run, uint32_t tag, float* out, int32_t count);
