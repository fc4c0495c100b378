/* FP64 CPU oracle of RGL's raytrace-and-point-cloud semantics -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and only as the
 * checker (or the timed CPU baseline). The product path never calls it.
 */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
	int nMeshes;
	const int32_t* meshVertOffset; /* nMeshes + 1 prefix offsets into vertices (in vertices) */
	const int32_t* meshTriOffset;  /* nMeshes + 1 prefix offsets into indices (in triangles) */
	const float* vertices;         /* xyz per vertex, all meshes concatenated */
	const int32_t* indices;        /* 3 mesh-local vertex indices per triangle */
	const float* uvs;              /* uv per vertex (same indexing as vertices) or NULL */
	const int32_t* meshHasUV;      /* per mesh */
	int nInstances;
	const int32_t* instMesh;
	const float* instTransform;     /* 12 floats (row-major 3x4) per instance */
	const float* instPrevTransform; /* 12 floats per instance (used when instHasPrev) */
	const int32_t* instHasPrev;
	const int32_t* instId;
	const float* instRetro;
	const int32_t* instTexture; /* texture index or -1 */
	int nTextures;
	const int32_t* texWidth;
	const int32_t* texHeight;
	const int64_t* texOffset; /* byte offset of each texture in texels */
	const uint8_t* texels;
	/* Vertex animation (src/scene/Entity.cpp:103-117): per instance, the float offset into `displacements` of its
	   object-space displacement since the previous frame (xyz per mesh vertex), or -1. NULL = none anywhere. */
	const int64_t* instDispOffset;
	const float* displacements;
} OrcScene;

typedef struct
{
	int nRays;
	const float* rays;          /* 12 floats per ray: world-frame rays after the rays chain */
	float rayOriginToWorld[12]; /* cumulative transform of the rays chain */
	const float* ranges;        /* 2 floats per entry */
	int nRanges;
	const int32_t* ringIds;
	int nRingIds;
	const float* timeOffsetsMs;
	const int8_t* mask;
	int distortion;
	float linVel[3];
	float angVel[3];
	float nearDistance;
	float farDistance;
	float defaultIntensity;
	float sceneDeltaTime;
	float hDivergence; /* full angles, radians (the API's values) */
	float vDivergence;
	int32_t returnMode;
	int bruteForce;
	int threads;
	double ambiguityEps; /* relative tolerance of the grazing / tie diagnostic */
	/* Optional: the same world-frame rays in FP64 (12 doubles per ray), e.g. the exact product of a sensor pose and
	   sensor-frame rays. When set they replace `rays`, and a ray's direction is its matrix's third column. */
	const double* raysF64;
} OrcRequest;

typedef struct
{
	float* xyz;
	int32_t* isHit;
	uint32_t* rayIdx;
	uint16_t* ringId;
	uint8_t* returnType;
	float* distance;
	float* intensityF32;
	uint8_t* intensityU8;
	float* laserRetro;
	double* timeStampF64;
	uint32_t* timeStampU32;
	int32_t* entityId;
	float* absVelocity;
	float* relVelocity;
	float* radialSpeed;
	float* azimuth;
	float* elevation;
	float* normal;
	float* incidentAngle;
	/* per beam diagnostics */
	int8_t* ambiguous;     /* bit 0: a sample grazes an edge, ties another instance or sits at a range boundary;
	                          bit 1: a distance comparison of the beam reduction is a near tie;
	                          bit 2: an intensity comparison of the beam reduction is a near tie;
	                          bit 3: a sample's texture coordinate lies at a texel edge;
	                          bit 4: the closest hit ties a hit of another instance (set with bit 0), e.g. coplanar
	                          faces of two entities */
	double* primaryT;      /* FP64 ray parameter of the primary sample's closest hit, -1 if none */
	int32_t* primaryInst;  /* instance of that hit, -1 if none */
	int32_t* primaryPrim;  /* triangle of that hit, -1 if none */
	/* per beam, 4 floats: the least and greatest SECOND distance and SECOND_STRONGEST intensity the beam reduction
	   can return when every comparison within FP32's reach of a tie may go either way (+-inf if unbounded: more
	   than 512 reachable tracker states, single-sample beams, no hit) */
	float* secondRange;
} OrcOutputs;

/* Runs the whole raytrace node (raygen, closest hit, miss, beam reduction) for every beam. Returns 0 on success. */
int orc_raytrace(const OrcScene* scene, const OrcRequest* request, OrcOutputs* out);

/* Persistent variant: builds the oracle's BVHs once; the scene arrays must stay alive until destroy. */
void* orc_world_create(const OrcScene* scene, int bruteForce);
int orc_world_raytrace(void* world, const OrcRequest* request, OrcOutputs* out);
void orc_world_destroy(void* world);

/* Plants one deliberate mistake in the oracle (0 = none), so tests can show that their checks catch it:
   1 = the beam reduction seeds its SECOND / SECOND_STRONGEST trackers with +inf / -inf instead of the first hit,
   2 = the LAST tracker keeps the minimum distance, 3 = velocity distortion moves the ray backwards in time,
   4 = texel lookup swaps the texture's x and y, 5 = azimuth reads the transposed matrix element,
   6 = the incidence factor uses the unnormalized normal. Process-wide; tests reset it to 0. */
void orc_plant_mistake(int which);

/* Philox4x32-10 block (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Standard normal variates of the library's seeded stream: normal(seed, run, tag, i), i in [0, count). */
void orc_normals(uint64_t seed, uint32_t run, uint32_t tag, float* out, int count);
float orc_normal_at(uint64_t seed, uint32_t run, uint32_t tag, uint32_t index);

/* ------------------------------------------------------------------------------------------------ CPU port
   FP32 CPU port of the library's lidar graph rays -> pose -> range -> raytrace -> points transform -> distance
   noise -> compact (cpu_port.cpp): the timed CPU baseline. Its own canonical LBVH (Morton codes, radix tree), built
   once per scene; frames run on `threads` std::threads over 64-ray tiles. */
typedef struct
{
	int nRays;
	const float* rays;    /* 12 floats per sensor-frame ray */
	float pose[12];       /* rays transform, composed with each ray in FP64 and rounded once */
	float range[2];
	float toSensor[12];   /* points transform */
	float lookAt[12];     /* look-at transform of the distance-noise node and its inverse */
	float lookAtInv[12];
	float noiseMean, noiseBase, noiseRise;
	uint64_t seed;
	uint32_t run, tag;
} CpuFrame;

typedef struct
{
	/* uncompacted outputs of the noise node (nRays entries each) */
	float* fullXyz;
	float* fullDistance;
	int32_t* fullEntityId;
	float* fullNormal;
	int32_t* fullIsHit;
	/* compacted hits (room for nRays); `kept` is written */
	float* xyz;
	float* distance;
	int32_t* entityId;
	float* normal;
	int32_t kept;
} CpuOut;

void* cpu_scene_create(const OrcScene* scene);
void cpu_scene_destroy(void* scene);
/* One frame of `count` sensors (sensor k on scenes[k]) on `threads` threads. Returns 0. */
int cpu_run(void* const* scenes, const CpuFrame* frames, CpuOut* outs, int count, int threads);

/* Natural log and sin/cos(2 pi u) with the single-precision methods the noise stream is defined with. */
float orc_logf(float x);
void orc_sincos2pi(float u, float* s, float* c);

#ifdef __cplusplus
}
#endif
