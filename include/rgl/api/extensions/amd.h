// RGL-MI355X extensions: entry points with no counterpart in the reference API.
//
// They expose what the MI355X implementation adds on top of RGL's C interface: control of the seeded noise
// stream (the reference seeds noise from std::random_device, src/graph/NodesCore.hpp:536,563,590, so it cannot
// be replayed), the TLAS update policy for moving entities, and introspection used by tests and benchmarks.
#pragma once

#include <rgl/api/core.h>

// Execution order of the graph containing `node`: its nodes in the order rgl_graph_run enqueues them, the reversed
// post-order of a depth-first search from the entry nodes taken in ascending priority (the order the reference's
// GraphRunCtx computes, src/graph/GraphRunCtx.cpp:107-142). On entry *inout_count is the capacity of out_nodes, on
// return the graph's node count; nodes beyond the capacity are not written. Nothing is run or validated.
RGL_API rgl_status_t rgl_graph_get_execution_order(rgl_node_t node, rgl_node_t* out_nodes, int32_t* inout_count);

// Sets the seed of a Gaussian noise node's stream and restarts its run counter. The noise drawn for element i in
// run r is a pure function of (seed, r, node kind, i), so the same seed replays the same noise on CPU and GPU.
// Radar postprocess nodes (rgl_node_points_radar_postprocess) draw their per-cluster noise from such a stream too
// and accept both calls.
RGL_API rgl_status_t rgl_node_gaussian_noise_set_seed(rgl_node_t node, uint64_t seed);

// Reports the stream identity (seed, next run index, node-kind tag) of a Gaussian noise node.
RGL_API rgl_status_t rgl_node_gaussian_noise_get_stream(rgl_node_t node, uint64_t* out_seed, uint32_t* out_run,
                                                        uint32_t* out_tag);

// Chooses how the scene's TLAS follows entity transform changes when no entity was added or removed:
// refit = true keeps the topology and refits boxes; false rebuilds the TLAS (default).
RGL_API rgl_status_t rgl_scene_set_tlas_refit(rgl_scene_t scene, bool refit);

// BVH node format used by traversal in every scene: 0 = binary with FP32 child boxes (64 B nodes), 1 = 4-wide with
// FP32 child boxes (128 B), 2 = 4-wide with child boxes quantized to 8 bits relative to the node box and rounded
// outward (64 B), 3 = the same with 16-bit bounds (80 B). All formats index the same leaves, so results are
// identical; they differ in speed only. Scenes are rebuilt on their next use.
RGL_API rgl_status_t rgl_configure_bvh_format(int32_t format);

// Reports the BVH node format in use.
RGL_API rgl_status_t rgl_get_bvh_format(int32_t* out_format);

// Per-ray traversal counters of a raytrace node's last configuration: internal BVH nodes visited, triangles tested
// and instances entered. Arrays hold `ray_count` entries (the node's ray count).
RGL_API rgl_status_t rgl_node_raytrace_traversal_stats(rgl_node_t node, int32_t* nodes_visited, int32_t* tris_tested,
                                                       int32_t* instances_entered, int32_t ray_count);

// Bytes of acceleration-structure and scene records the raytrace launches of the given raytrace nodes read, each
// 64 B line counted once over all of them (the unique, compulsory part of their BVH traffic): nodes in the node format
// in use, leaf triangles, instance records, and the shading records, index triples, vertices and texture coordinates
// of the triangles hit. Traces the nodes' current rays again (a measurement; results are not touched).
RGL_API rgl_status_t rgl_graph_traversal_footprint(const rgl_node_t* raytrace_nodes, int32_t count, int64_t* out_bytes);

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

// Point counts of `field` at nodes[0..node_count) (rgl_graph_get_result_size of each node, in one call; sizes[i]
// receives the count of nodes[i]). A batch's counts arrive on the host together, so reading them for many sensors
// costs one wait and one call instead of one per node; the first failing node fails the call.
RGL_API rgl_status_t rgl_graph_get_result_sizes(const rgl_node_t* nodes, int32_t node_count, rgl_field_t field,
                                                int32_t* out_counts);

// Autotuning of rgl_graph_run_batch's launch shapes (on by default): while untuned, successive batches alternate
// between the raytrace block sizes (128 or 64 threads) and the compaction tiles (8 or 4 points per thread), each
// launch of at least 100k rays / points timed with HIP events; after three timings per variant the fastest per
// element is kept. Variants never change results. configure_autotune(true) restarts tuning; force_variants fixes
// the variants and disables tuning (raytrace 0, 1, or 2 = a persistent grid of 64-thread blocks claiming runs of
// blocks from a counter, which the tuner does not consider: it measured slower than hardware block dispatch;
// compaction 0 or 1); get_variants reports the variants in use (-1 while tuning).
RGL_API rgl_status_t rgl_graph_batch_configure_autotune(bool enable);
RGL_API rgl_status_t rgl_graph_batch_force_variants(int32_t raytrace_variant, int32_t compact_variant);
RGL_API rgl_status_t rgl_graph_batch_get_variants(int32_t* out_raytrace_variant, int32_t* out_compact_variant);

// HIP graph capture of rgl_graph_run_batch (off by default; worth enabling when batches are small and launch-bound): a batch whose kernel sequence (item counts per stage,
// grid shapes, launch variants, descriptor buffers) equals the previous batch's is captured into a HIP graph, and
// later batches with that same sequence replay it in one launch; descriptor contents (node buffers, scene
// pointers, parameters) may change freely between replays. Batches that are being timed or tuned, or that contain
// nodes without a batched form, run launch by launch. get_capture_stats reports captures and replays so far.
RGL_API rgl_status_t rgl_graph_batch_configure_capture(bool enable);
RGL_API rgl_status_t rgl_graph_batch_get_capture_stats(int64_t* out_captures, int64_t* out_replays);

// Number of elementwise point stages (points transform, distance noise) computed inside a batched raytrace launch
// instead of a launch of their own, since the library was loaded. rgl_graph_run_batch folds into the launch of a
// single-sample raytrace node the chain of such nodes that starts at its only output and continues through nodes
// with one output (at most two stages); each folded node still writes its own results.
RGL_API rgl_status_t rgl_graph_batch_get_fused_stages(int64_t* out_stages);

// Enables (and resets) HIP-event timing of the batched raytrace launches.
RGL_API rgl_status_t rgl_graph_batch_set_kernel_timing(bool enable);

// Total device time (ms) and number of batched raytrace launches timed since timing was enabled; waits for them.
RGL_API rgl_status_t rgl_graph_batch_get_kernel_timing(double* out_total_ms, int32_t* out_launches);

// Prefetch of a yield node's results (the reference prefetches XYZ of every yield node to pinned host memory,
// src/graph/YieldPointsNode.cpp:31-35): when enabled, every run of the graph also copies the live points of every
// yielded field into pinned host memory with a kernel that writes over the bus, so the results are host-resident
// when the run completes and rgl_graph_get_result_data is a host copy. Off by default.
RGL_API rgl_status_t rgl_node_points_yield_configure_prefetch(rgl_node_t node, bool enable);

// Radar range x azimuth binning (build-defined; the reference instead clusters radar hits by scopes with
// physical-optics RCS, src/graph/RadarPostprocessPointsNode.cpp:88-189). Input: XYZ, DISTANCE, INCIDENT_ANGLE and
// IS_HIT. A hit at distance d whose point has sensor-frame azimuth a = atan2(y, x) (sensor frame = the node's
// look-at origin, i.e. before the rays' cumulative transform) falls in cell (floor(d / range_step),
// floor((a - azimuth_min) / azimuth_step)); hits outside [0, range_max) x [azimuth_min, azimuth_max) are dropped.
// Its RCS contribution is d^2 * ray_solid_angle * cos(incident angle) m^2 (the ray's footprint area projected
// back along the ray), rounded to 2^-24 m^2 and summed exactly per cell. Output: one point per non-empty cell, in
// cell order (range-major): XYZ (cell centre in the sensor frame, z = 0), DISTANCE and AZIMUTH (cell centre),
// RCS_F32 (10 log10 of the summed contributions, dBsm), IS_HIT (1).
RGL_API rgl_status_t rgl_node_points_radar_bin(rgl_node_t* node, float range_step, float range_max, float azimuth_step,
                                               float azimuth_min, float azimuth_max, float ray_solid_angle);

// ---------------------------------------------------------------------------------------------------- serving
// A request server doing continuous batching of sensor frames (one graph per sensor) on the calling thread.
// Requests queue in FIFO order; rgl_server_step retires finished batches and launches the admissible queued requests
// into the free slots, spread evenly over them (at most `max_batch_graphs` per batch; rgl_graph_run_batch semantics),
// so up to `slots` batches overlap on the device and the requests of a finished batch run again while the others
// still occupy it. A request is admissible when its graph has no earlier request in flight and the acceleration structures
// of the scenes it raytraces fit `hbm_budget_bytes` (<= 0: unlimited) after evicting least recently used idle
// scenes (evicted scenes rebuild on their next use); a request whose scenes alone exceed the budget fails (state 3).
// With `batching` false every request runs alone through
// rgl_graph_run semantics (one per slot): the plain per-sensor path under the same queueing.
// A request's results are read with rgl_graph_get_result_* once it is done, until its graph runs again.
typedef struct RglServer* rgl_server_t;

typedef struct
{
	int64_t submitted;
	int64_t completed;
	int64_t batches;
	int64_t rays_completed;  // rays of completed requests (sum over their rays nodes)
	int64_t evictions;       // scenes whose acceleration structures were evicted
	int64_t deferred;        // times a queued request had to wait for its graph or for scene memory
	int64_t failed;          // requests whose scenes alone exceed the HBM budget (state 3)
	int64_t resident_bytes;  // acceleration-structure bytes of the admitted scenes
	double p50_ms, p99_ms, mean_ms, max_ms; // request latency, submit to retirement
} rgl_server_stats_t;

RGL_API rgl_status_t rgl_server_create(rgl_server_t* out_server, int32_t slots, int32_t max_batch_graphs,
                                       int64_t hbm_budget_bytes, bool batching);
RGL_API rgl_status_t rgl_server_destroy(rgl_server_t server); // waits for queued and running requests
// Queues one frame of the graph containing `node`; returns its ticket.
RGL_API rgl_status_t rgl_server_submit(rgl_server_t server, rgl_node_t node, int64_t* out_ticket);
// Retires finished batches and launches queued requests into free slots.
RGL_API rgl_status_t rgl_server_step(rgl_server_t server, int32_t* out_retired);
// State of a ticket: 0 queued, 1 running, 2 done, 3 failed; latency in ms once done.
RGL_API rgl_status_t rgl_server_poll(rgl_server_t server, int64_t ticket, int32_t* out_state, double* out_latency_ms);
// Steps until the ticket is done.
RGL_API rgl_status_t rgl_server_wait(rgl_server_t server, int64_t ticket, double* out_latency_ms);
RGL_API rgl_status_t rgl_server_get_stats(rgl_server_t server, rgl_server_stats_t* out_stats);
RGL_API rgl_status_t rgl_server_reset_stats(rgl_server_t server);

// Standard normal variates normal(seed, run, tag, i) for i in [0, count), computed on the GPU.
RGL_API rgl_status_t rgl_noise_normals(uint64_t seed, uint32_t run, uint32_t tag, float* out, int32_t count);

// rgl_noise_normals without any device allocation. No counterpart in the reference: it is for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
// variates are staged in a fixed per-device array of 2^20 floats (4 MiB on every device the library runs on, held
// until the process ends), so count must not exceed 1048576 (RGL_INVALID_ARGUMENT otherwise). Synchronous; calls that
// overlap on one GPU share the staging array and race.
RGL_API rgl_status_t rgl_noise_normals_no_alloc(uint64_t seed, uint32_t run, uint32_t tag, float* out, int32_t count);
