// extern "C" entry points of the RGL API (include/rgl/api/core.h), implemented over the HIP runtime.
//
// Conventions follow the reference's API layer (src/api/apiCommon.hpp:39-139, src/api/apiCommon.cpp:44-90,
// src/api/apiCore.cpp): every call runs inside safeCall(), which performs lazy initialization, refuses work after
// an unrecoverable status, and maps exceptions to rgl_status_t; argument checks produce
// "RGL API Error: Invalid argument, condition unsatisfied: <expr>"; nodes are created or updated in place.

#include <rgl/api/core.h>
#include <rgl/api/extensions/amd.h>
#include <rgl/api/extensions/tape.h>

#include <cstdio>
#include <fstream>
#include <functional>
#include <memory>
#include <mutex>
#include <optional>
#include <set>
#include <string>

#include <cstring>

#include "graph.hpp"
#include "hostcopy.hpp"
#include "server.hpp"
#include "tape.hpp"
#include "scene.hpp"

using namespace rgl;

namespace {

rgl_status_t lastStatus = RGL_SUCCESS;
std::optional<std::string> lastMessage;
int logLevel = RGL_LOG_LEVEL_INFO;
bool logStdout = true;
std::unique_ptr<std::ofstream> logFile;

struct NotImplementedCall : std::runtime_error
{
	using std::runtime_error::runtime_error;
};

const char* fallbackMessage(rgl_status_t s)
{
	switch (s) {
		case RGL_SUCCESS: return "operation successful";
		case RGL_INVALID_ARGUMENT: return "invalid argument";
		case RGL_INTERNAL_EXCEPTION: return "unhandled internal exception";
		case RGL_INVALID_STATE: return "invalid state - unrecoverable error occurred";
		case RGL_NOT_IMPLEMENTED: return "operation not (yet) implemented";
		case RGL_LOGGING_ERROR: return "logging error";
		case RGL_INVALID_FILE_PATH: return "invalid file path";
		case RGL_TAPE_ERROR: return "tape error";
		case RGL_UDP_ERROR: return "UDP error";
		case RGL_ROS2_ERROR: return "ROS2 error";
		default: return "???";
	}
}

bool recoverable(rgl_status_t s)
{
	return s == RGL_SUCCESS || s == RGL_INVALID_ARGUMENT || s == RGL_INVALID_API_OBJECT || s == RGL_INVALID_PIPELINE ||
	       s == RGL_INVALID_FILE_PATH || s == RGL_NOT_IMPLEMENTED || s == RGL_TAPE_ERROR || s == RGL_UDP_ERROR ||
	       s == RGL_ROS2_ERROR;
}

void log(int level, const std::string& msg)
{
	if (level < logLevel || logLevel >= RGL_LOG_LEVEL_OFF) {
		return;
	}
	static const char* names[] = {"trace", "debug", "info", "warn", "error", "critical"};
	const std::string line = std::string("[RGL ") + names[std::min(level, 5)] + "] " + msg + "\n";
	if (logStdout) {
		std::fputs(line.c_str(), stdout);
		std::fflush(stdout);
	}
	if (logFile) {
		*logFile << line;
		logFile->flush();
	}
}

rgl_status_t updateState(rgl_status_t s, std::optional<std::string> msg = std::nullopt)
{
	lastStatus = s;
	lastMessage = std::move(msg);
	if (s != RGL_SUCCESS) {
		const std::string text = lastMessage.value_or(fallbackMessage(s));
		log(recoverable(s) ? RGL_LOG_LEVEL_ERROR : RGL_LOG_LEVEL_CRITICAL,
		    (recoverable(s) ? "Recoverable error (code=" : "Unrecoverable error (code=") + std::to_string(s) + "): " + text);
	}
	return s;
}

template<typename Fn>
rgl_status_t safeCall(Fn&& fn);

void lazyInit()
{
	static bool called = false;
	if (called) {
		return;
	}
	called = true;
	rgl_status_t st = safeCall([]() {
		int count = 0;
		RGL_HIP(hipGetDeviceCount(&count));
		if (count <= 0) {
			throw std::runtime_error("no HIP device available");
		}
		RGL_HIP(hipSetDevice(0));
		hipDeviceProp_t prop{};
		RGL_HIP(hipGetDeviceProperties(&prop, 0));
		log(RGL_LOG_LEVEL_DEBUG, std::string("RGL-MI355X on ") + prop.name + " (" + prop.gcnArchName + ")");
	});
	if (st != RGL_SUCCESS) {
		updateState(RGL_INITIALIZATION_ERROR, lastMessage);
	}
}

template<typename Fn>
rgl_status_t safeCall(Fn&& fn)
{
	lazyInit();
	if (!recoverable(lastStatus)) {
		return updateState(RGL_INVALID_STATE);
	}
	try {
		fn();
	} catch (NotImplementedCall& e) {
		return updateState(RGL_NOT_IMPLEMENTED, e.what());
	} catch (tape::TapeError& e) {
		return updateState(RGL_TAPE_ERROR, e.what());
	} catch (tape::InvalidFilePath& e) {
		return updateState(RGL_INVALID_FILE_PATH, e.what());
	} catch (InvalidAPIObject& e) {
		return updateState(RGL_INVALID_API_OBJECT, e.what());
	} catch (InvalidPipeline& e) {
		return updateState(RGL_INVALID_PIPELINE, e.what());
	} catch (std::invalid_argument& e) {
		return updateState(RGL_INVALID_ARGUMENT, e.what());
	} catch (std::exception& e) {
		return updateState(RGL_INTERNAL_EXCEPTION, e.what());
	} catch (...) {
		return updateState(RGL_INTERNAL_EXCEPTION, "exceptional exception");
	}
	return updateState(RGL_SUCCESS);
}

#define CHECK_ARG(expr)                                                                                                \
	do {                                                                                                               \
		if (!(expr)) {                                                                                                 \
			throw std::invalid_argument(std::string("RGL API Error: Invalid argument, condition unsatisfied: ") + #expr); \
		}                                                                                                              \
	} while (0)

rgl_node_t handleOf(const rgl::Node::Ptr& n) { return reinterpret_cast<rgl_node_t>(n.get()); }

template<typename T>
std::shared_ptr<T> nodeAs(rgl_node_t h, const char* expected)
{
	auto base = lookupNode(h);
	auto typed = std::dynamic_pointer_cast<T>(base);
	if (!typed) {
		throw InvalidAPIObject(std::string("RGL API Error: Node type mismatch: expected ") + expected + ", got " +
		                       base->getName());
	}
	return typed;
}

// src/api/apiCommon.hpp:108-139
template<typename T, typename Fn>
void createOrUpdateNode(rgl_node_t* handle, const char* typeName, bool graphValidationNeeded, Fn&& setParams)
{
	std::shared_ptr<T> node;
	if (*handle == nullptr) {
		node = std::make_shared<T>();
	} else {
		node = nodeAs<T>(*handle, typeName);
	}
	if (node->hasCtx()) {
		node->getCtx()->synchronize();
	}
	if (graphValidationNeeded && node->hasCtx()) {
		node->getCtx()->markNodesDirty();
	}
	setParams(*node);
	node->markDirty();
	if (*handle == nullptr) {
		registerNode(node);
	}
	*handle = handleOf(node);
}

void syncNodeGraph(const rgl::Node::Ptr& n)
{
	if (n->hasCtx()) {
		n->getCtx()->synchronize();
	}
}

void checkOverflow(const rgl::Node::Ptr& n)
{
	if (!n->hasCtx()) {
		return;
	}
	for (auto& m : n->getCtx()->getNodes()) {
		if (auto rt = std::dynamic_pointer_cast<RaytraceNode>(m)) {
			if (rt->overflowed()) {
				throw std::runtime_error("BVH traversal stack overflow: scene hierarchy is deeper than supported");
			}
		}
	}
}

std::shared_ptr<rgl::Scene> sceneOf(rgl_scene_t scene)
{
	return scene == nullptr ? rgl::Scene::defaultShared() : Registry<rgl::Scene>::get(scene, "Scene");
}

Mat3x4f toMat(const rgl_mat3x4f* m)
{
	Mat3x4f r;
	std::memcpy(r.rc, m->value, sizeof(r.rc));
	return r;
}

} // namespace

namespace rgl {
std::string fieldName(rgl_field_t f)
{
	switch (f) {
		case RGL_FIELD_XYZ_VEC3_F32: return "XYZ_VEC3_F32";
		case RGL_FIELD_INTENSITY_F32: return "INTENSITY_F32";
		case RGL_FIELD_INTENSITY_U8: return "INTENSITY_U8";
		case RGL_FIELD_IS_HIT_I32: return "IS_HIT_I32";
		case RGL_FIELD_IS_GROUND_I32: return "IS_GROUND_I32";
		case RGL_FIELD_RAY_IDX_U32: return "RAY_IDX_U32";
		case RGL_FIELD_ENTITY_ID_I32: return "ENTITY_ID_I32";
		case RGL_FIELD_DISTANCE_F32: return "DISTANCE_F32";
		case RGL_FIELD_AZIMUTH_F32: return "AZIMUTH_F32";
		case RGL_FIELD_ELEVATION_F32: return "ELEVATION_F32";
		case RGL_FIELD_RING_ID_U16: return "RING_ID_U16";
		case RGL_FIELD_RETURN_TYPE_U8: return "RETURN_TYPE_U8";
		case RGL_FIELD_TIME_STAMP_F64: return "TIME_STAMP_F64";
		case RGL_FIELD_TIME_STAMP_U32: return "TIME_STAMP_U32";
		case RGL_FIELD_ABSOLUTE_VELOCITY_VEC3_F32: return "ABSOLUTE_VELOCITY_VEC3_F32";
		case RGL_FIELD_RELATIVE_VELOCITY_VEC3_F32: return "RELATIVE_VELOCITY_VEC3_F32";
		case RGL_FIELD_RADIAL_SPEED_F32: return "RADIAL_SPEED_F32";
		case RGL_FIELD_POWER_F32: return "POWER_F32";
		case RGL_FIELD_RCS_F32: return "RCS_F32";
		case RGL_FIELD_NOISE_F32: return "NOISE_F32";
		case RGL_FIELD_SNR_F32: return "SNR_F32";
		case RGL_FIELD_NORMAL_VEC3_F32: return "NORMAL_VEC3_F32";
		case RGL_FIELD_INCIDENT_ANGLE_F32: return "INCIDENT_ANGLE_F32";
		case RGL_FIELD_RAY_POSE_MAT3x4_F32: return "RAY_POSE_MAT3x4_F32";
		case RGL_FIELD_LASER_RETRO_F32: return "LASER_RETRO_F32";
		case RGL_FIELD_PADDING_8: return "PADDING_8";
		case RGL_FIELD_PADDING_16: return "PADDING_16";
		case RGL_FIELD_PADDING_32: return "PADDING_32";
		case RGL_FIELD_DYNAMIC_FORMAT: return "DYNAMIC_FORMAT";
		default: return "UNKNOWN(" + std::to_string(static_cast<int>(f)) + ")";
	}
}

void throwHip(hipError_t e, const char* expr, const char* file, int line)
{
	throw HipError(std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ") at " + file + ":" +
	               std::to_string(line) + ": " + expr);
}
} // namespace rgl

// ============================================================================================ general
// ============================================================================================ tape recording
// Arguments of a recorded call as YAML scalars (src/tape/TapeRecorder.hpp:69-128): handles by address, out-handles
// by the handle written, int32_t out-parameters by value, enums as integers, arrays and single structs by their
// offset in the binary file.
namespace tapeargs {
using rgl::tape::handle;
using rgl::tape::scalar;
inline std::string A(int32_t v) { return scalar(int64_t{v}); }
inline std::string A(uint64_t v) { return rgl::tape::scalarU(v); }
inline std::string A(float v) { return scalar(static_cast<double>(v), true); }
inline std::string A(bool v) { return scalar(v); }
inline std::string A(const char* s) { return scalar(s); }
inline std::string A(const void* p) { return handle(p); }
inline std::string A(rgl_mesh_t h) { return handle(h); }
inline std::string A(rgl_entity_t h) { return handle(h); }
inline std::string A(rgl_texture_t h) { return handle(h); }
inline std::string A(rgl_node_t h) { return handle(h); }
inline std::string A(rgl_scene_t h) { return handle(h); }
template<typename E>
std::enable_if_t<std::is_enum_v<E>, std::string> A(E e)
{
	return scalar(static_cast<int64_t>(e));
}
template<typename H>
std::string D(H* out)
{
	return out != nullptr ? A(*out) : std::string("0");
}
inline std::string V(const int32_t* p) { return p != nullptr ? A(*p) : std::string("0"); }
template<typename T>
std::string R(const T* p, int64_t count)
{
	return rgl::tape::active()->bytes(rgl::tape::array(p, count));
}
} // namespace tapeargs

// Bytes of one record of `fields` (0 if any field is unknown).
int64_t recordBytes(const rgl_field_t* fields, int32_t count)
{
	int64_t b = 0;
	for (int32_t i = 0; fields != nullptr && i < count; ++i) {
		try {
			b += static_cast<int64_t>(fieldSizeChecked(fields[i]));
		} catch (...) {
			return 0;
		}
	}
	return b;
}

template<typename MakeArgs>
void tapeRecord(rgl_status_t& status, const char* fn, MakeArgs makeArgs)
{
	auto* rec = tape::active();
	if (rec == nullptr) {
		return;
	}
	try {
		rec->call(fn, makeArgs());
	} catch (std::exception& e) {
		status = updateState(RGL_TAPE_ERROR, e.what());
	}
}

#define TAPE_RECORD(fn, ...)                                                                                           \
	tapeRecord(status, #fn, [&]() {                                                                                    \
		using namespace tapeargs;                                                                                      \
		return std::vector<std::string>{__VA_ARGS__};                                                                  \
	})

RGL_API rgl_status_t rgl_get_version_info(int32_t* out_major, int32_t* out_minor, int32_t* out_patch)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(out_major != nullptr);
		CHECK_ARG(out_minor != nullptr);
		CHECK_ARG(out_patch != nullptr);
		*out_major = RGL_VERSION_MAJOR;
		*out_minor = RGL_VERSION_MINOR;
		*out_patch = RGL_VERSION_PATCH;
	});
	TAPE_RECORD(rgl_get_version_info, V(out_major), V(out_minor), V(out_patch));
	return status;
}

RGL_API rgl_status_t rgl_get_extension_info(rgl_extension_t extension, int32_t* out_available)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(0 <= extension && extension < RGL_EXTENSION_COUNT);
		CHECK_ARG(out_available != nullptr);
		*out_available = 0;
	});
	TAPE_RECORD(rgl_get_extension_info, A(extension), V(out_available));
	return status;
}

RGL_API rgl_status_t rgl_configure_logging(rgl_log_level_t log_level, const char* log_file_path, bool use_stdout)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(0 <= log_level && log_level <= 6);
		logLevel = log_level;
		logStdout = use_stdout;
		logFile.reset();
		if (log_file_path != nullptr) {
			logFile = std::make_unique<std::ofstream>(log_file_path, std::ios::trunc);
			if (!logFile->good()) {
				logFile.reset();
				throw std::runtime_error(std::string("cannot open log file ") + log_file_path);
			}
		}
	});
	TAPE_RECORD(rgl_configure_logging, A(log_level), A(log_file_path), A(use_stdout));
	return status;
}

RGL_API void rgl_get_last_error_string(const char** out_error_string)
{
	if (out_error_string == nullptr) {
		return;
	}
	*out_error_string = lastMessage.has_value() ? lastMessage->c_str() : fallbackMessage(lastStatus);
}

RGL_API rgl_status_t rgl_cleanup(void)
{
	rgl_status_t status = safeCall([&]() {
		for (auto& [h, sv] : Registry<rgl::Server>::map()) {
			sv->drain();
		}
		Registry<rgl::Server>::map().clear();
		GraphBatch::instance().synchronize();
		for (auto& n : allNodes()) {
			if (n->hasCtx()) {
				n->getCtx()->detachAndDestroy();
			}
		}
		for (auto& n : allNodes()) {
			n->disconnectConnectedNodes();
			releaseNode(n.get());
		}
		Registry<rgl::Entity>::map().clear();
		Registry<rgl::Mesh>::map().clear();
		Registry<rgl::Texture>::map().clear();
		for (auto& [h, sc] : Registry<rgl::Scene>::map()) {
			sc->clear();
		}
		Registry<rgl::Scene>::map().clear();
		rgl::Scene::instance().clear();
	});
	TAPE_RECORD(rgl_cleanup);
	return status;
}

// ============================================================================================ mesh
RGL_API rgl_status_t rgl_mesh_create(rgl_mesh_t* out_mesh, const rgl_vec3f* vertices, int32_t vertex_count,
                                     const rgl_vec3i* indices, int32_t index_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(out_mesh != nullptr);
		CHECK_ARG(vertices != nullptr);
		CHECK_ARG(vertex_count > 0);
		CHECK_ARG(indices != nullptr);
		CHECK_ARG(index_count > 0);
		auto m = std::make_shared<rgl::Mesh>(reinterpret_cast<const Vec3f*>(vertices), vertex_count,
		                                reinterpret_cast<const Vec3i*>(indices), index_count);
		Registry<rgl::Mesh>::add(m);
		*out_mesh = reinterpret_cast<rgl_mesh_t>(m.get());
	});
	TAPE_RECORD(rgl_mesh_create, D(out_mesh), R(vertices, vertex_count), A(vertex_count), R(indices, index_count), A(index_count));
	return status;
}

RGL_API rgl_status_t rgl_mesh_set_texture_coords(rgl_mesh_t mesh, const rgl_vec2f* uvs, int32_t uv_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(mesh != nullptr);
		CHECK_ARG(uvs != nullptr);
		CHECK_ARG(uv_count > 0);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Mesh>::get(mesh, "Mesh")->setTexCoords(reinterpret_cast<const Vec2f*>(uvs), uv_count);
		rgl::Scene::markAllDirty(false);
	});
	TAPE_RECORD(rgl_mesh_set_texture_coords, A(mesh), R(uvs, uv_count), A(uv_count));
	return status;
}

RGL_API rgl_status_t rgl_mesh_set_bone_weights(rgl_mesh_t mesh, const rgl_bone_weights_t* bone_weights,
                                               int32_t bone_weights_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(mesh != nullptr);
		CHECK_ARG(bone_weights != nullptr);
		CHECK_ARG(bone_weights_count > 0);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Mesh>::get(mesh, "Mesh")->setBoneWeights(reinterpret_cast<const BoneWeights*>(bone_weights),
		                                                      bone_weights_count);
	});
	TAPE_RECORD(rgl_mesh_set_bone_weights, A(mesh), R(bone_weights, bone_weights_count), A(bone_weights_count));
	return status;
}

RGL_API rgl_status_t rgl_mesh_set_restposes(rgl_mesh_t mesh, const rgl_mat3x4f* restposes, int32_t bones_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(mesh != nullptr);
		CHECK_ARG(restposes != nullptr);
		CHECK_ARG(bones_count > 0);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Mesh>::get(mesh, "Mesh")->setRestposes(reinterpret_cast<const Mat3x4f*>(restposes), bones_count);
	});
	TAPE_RECORD(rgl_mesh_set_restposes, A(mesh), R(restposes, bones_count), A(bones_count));
	return status;
}

RGL_API rgl_status_t rgl_mesh_destroy(rgl_mesh_t mesh)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(mesh != nullptr);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Mesh>::release(mesh, "Mesh");
	});
	TAPE_RECORD(rgl_mesh_destroy, A(mesh));
	return status;
}

RGL_API rgl_status_t rgl_mesh_is_alive(rgl_mesh_t mesh, bool* out_alive)
{
	return safeCall([&]() {
		CHECK_ARG(out_alive != nullptr);
		*out_alive = Registry<rgl::Mesh>::contains(mesh);
	});
}

// ============================================================================================ entity
// Calls that change only host state of a scene (entity transforms, ids, laser retro, textures, scene time, new
// entities) return without waiting for running graphs, where src/api/apiCore.cpp:333-425 drains every graph: the
// scene's next rebuild runs on the scene stream after the raytrace launches that still read its device arrays
// (Scene::rebuild). Calls that rewrite device data those launches read (animation, texture coordinates) or free it
// still drain every graph first.
RGL_API rgl_status_t rgl_entity_create(rgl_entity_t* out_entity, rgl_scene_t scene, rgl_mesh_t mesh)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(out_entity != nullptr);
		CHECK_ARG(mesh != nullptr);
		auto sc = sceneOf(scene);
		auto e = std::make_shared<rgl::Entity>(Registry<rgl::Mesh>::get(mesh, "Mesh"), sc.get());
		Registry<rgl::Entity>::add(e);
		sc->addEntity(e);
		*out_entity = reinterpret_cast<rgl_entity_t>(e.get());
	});
	TAPE_RECORD(rgl_entity_create, D(out_entity), A(scene), A(mesh));
	return status;
}

RGL_API rgl_status_t rgl_entity_destroy(rgl_entity_t entity)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		GraphRunCtx::synchronizeAll();
		auto e = Registry<rgl::Entity>::get(entity, "Entity");
		e->scene->removeEntity(e);
		Registry<rgl::Entity>::release(entity, "Entity");
	});
	TAPE_RECORD(rgl_entity_destroy, A(entity));
	return status;
}

RGL_API rgl_status_t rgl_entity_set_transform(rgl_entity_t entity, const rgl_mat3x4f* transform)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		CHECK_ARG(transform != nullptr);
		Registry<rgl::Entity>::get(entity, "Entity")->setTransform(toMat(transform));
	});
	TAPE_RECORD(rgl_entity_set_transform, A(entity), R(transform, 1));
	return status;
}

RGL_API rgl_status_t rgl_entity_set_pose_world(rgl_entity_t entity, const rgl_mat3x4f* pose, int32_t bones_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		CHECK_ARG(pose != nullptr);
		CHECK_ARG(bones_count > 0);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Entity>::get(entity, "Entity")->setPoseAndAnimate(reinterpret_cast<const Mat3x4f*>(pose), bones_count);
	});
	TAPE_RECORD(rgl_entity_set_pose_world, A(entity), R(pose, bones_count), A(bones_count));
	return status;
}

RGL_API rgl_status_t rgl_entity_set_id(rgl_entity_t entity, int32_t id)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		CHECK_ARG(id != RGL_ENTITY_INVALID_ID);
		auto e = Registry<rgl::Entity>::get(entity, "Entity");
		constexpr int maxId = 1 << 28; // entity ids are 28-bit (src/scene/Entity.cpp:45-48)
		if (id >= maxId) {
			throw std::invalid_argument("Entity ID (" + std::to_string(id) + ") must be less than " + std::to_string(maxId));
		}
		e->id = id;
		e->scene->markDirty(false);
	});
	TAPE_RECORD(rgl_entity_set_id, A(entity), A(id));
	return status;
}

RGL_API rgl_status_t rgl_entity_set_intensity_texture(rgl_entity_t entity, rgl_texture_t texture)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		CHECK_ARG(texture != nullptr);
		auto e = Registry<rgl::Entity>::get(entity, "Entity");
		e->texture = Registry<rgl::Texture>::get(texture, "Texture");
		e->scene->markDirty(false);
	});
	TAPE_RECORD(rgl_entity_set_intensity_texture, A(entity), A(texture));
	return status;
}

RGL_API rgl_status_t rgl_entity_set_laser_retro(rgl_entity_t entity, float retro)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		auto e = Registry<rgl::Entity>::get(entity, "Entity");
		e->laserRetro = retro;
		e->scene->markDirty(false);
	});
	TAPE_RECORD(rgl_entity_set_laser_retro, A(entity), A(retro));
	return status;
}

RGL_API rgl_status_t rgl_entity_apply_external_animation(rgl_entity_t entity, const rgl_vec3f* vertices,
                                                         int32_t vertex_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(entity != nullptr);
		CHECK_ARG(vertices != nullptr);
		CHECK_ARG(vertex_count > 0);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Entity>::get(entity, "Entity")->applyExternalAnimation(reinterpret_cast<const Vec3f*>(vertices),
		                                                                    vertex_count);
	});
	TAPE_RECORD(rgl_entity_apply_external_animation, A(entity), R(vertices, vertex_count), A(vertex_count));
	return status;
}

RGL_API rgl_status_t rgl_entity_is_alive(rgl_entity_t entity, bool* out_alive)
{
	return safeCall([&]() {
		CHECK_ARG(out_alive != nullptr);
		*out_alive = Registry<rgl::Entity>::contains(entity);
	});
}

// ============================================================================================ texture
RGL_API rgl_status_t rgl_texture_create(rgl_texture_t* out_texture, const void* texels, int32_t width, int32_t height)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(out_texture != nullptr);
		CHECK_ARG(texels != nullptr);
		CHECK_ARG(width > 0);
		CHECK_ARG(height > 0);
		auto t = std::make_shared<rgl::Texture>(texels, width, height);
		Registry<rgl::Texture>::add(t);
		*out_texture = reinterpret_cast<rgl_texture_t>(t.get());
	});
	TAPE_RECORD(rgl_texture_create, D(out_texture), R(static_cast<const uint8_t*>(texels), static_cast<int64_t>(width) * height), A(width), A(height));
	return status;
}

RGL_API rgl_status_t rgl_texture_destroy(rgl_texture_t texture)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(texture != nullptr);
		GraphRunCtx::synchronizeAll();
		Registry<rgl::Texture>::release(texture, "Texture");
	});
	TAPE_RECORD(rgl_texture_destroy, A(texture));
	return status;
}

RGL_API rgl_status_t rgl_texture_is_alive(rgl_texture_t texture, bool* out_alive)
{
	return safeCall([&]() {
		CHECK_ARG(out_alive != nullptr);
		*out_alive = Registry<rgl::Texture>::contains(texture);
	});
}

// ============================================================================================ scene
RGL_API rgl_status_t rgl_scene_set_time(rgl_scene_t scene, uint64_t nanoseconds)
{
	rgl_status_t status = safeCall([&]() {
		auto sc = sceneOf(scene);
		sc->setTime(nanoseconds);
		sc->markDirty(false);
	});
	TAPE_RECORD(rgl_scene_set_time, A(scene), A(nanoseconds));
	return status;
}

// ============================================================================================ rays nodes
RGL_API rgl_status_t rgl_node_rays_from_mat3x4f(rgl_node_t* node, const rgl_mat3x4f* rays, int32_t ray_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(rays != nullptr);
		CHECK_ARG(ray_count > 0);
		createOrUpdateNode<FromMat3x4fRaysNode>(node, "FromMat3x4fRaysNode", true, [&](FromMat3x4fRaysNode& n) {
			n.setParameters(reinterpret_cast<const Mat3x4f*>(rays), ray_count);
		});
	});
	TAPE_RECORD(rgl_node_rays_from_mat3x4f, D(node), R(rays, ray_count), A(ray_count));
	return status;
}

RGL_API rgl_status_t rgl_node_rays_set_ring_ids(rgl_node_t* node, const int32_t* ring_ids, int32_t ring_ids_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(ring_ids != nullptr);
		CHECK_ARG(ring_ids_count > 0);
		createOrUpdateNode<SetRingIdsRaysNode>(node, "SetRingIdsRaysNode", false,
		                                       [&](SetRingIdsRaysNode& n) { n.setParameters(ring_ids, ring_ids_count); });
	});
	TAPE_RECORD(rgl_node_rays_set_ring_ids, D(node), R(ring_ids, ring_ids_count), A(ring_ids_count));
	return status;
}

RGL_API rgl_status_t rgl_node_rays_set_range(rgl_node_t* node, const rgl_vec2f* ranges, int32_t ranges_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(ranges != nullptr);
		CHECK_ARG(ranges_count > 0);
		createOrUpdateNode<SetRangeRaysNode>(node, "SetRangeRaysNode", false, [&](SetRangeRaysNode& n) {
			n.setParameters(reinterpret_cast<const Vec2f*>(ranges), ranges_count);
		});
	});
	TAPE_RECORD(rgl_node_rays_set_range, D(node), R(ranges, ranges_count), A(ranges_count));
	return status;
}

RGL_API rgl_status_t rgl_node_rays_set_time_offsets(rgl_node_t* node, const float* offsets, int32_t offsets_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(offsets != nullptr);
		CHECK_ARG(offsets_count > 0);
		createOrUpdateNode<SetTimeOffsetsRaysNode>(node, "SetTimeOffsetsRaysNode", false,
		                                           [&](SetTimeOffsetsRaysNode& n) { n.setParameters(offsets, offsets_count); });
	});
	TAPE_RECORD(rgl_node_rays_set_time_offsets, D(node), R(offsets, offsets_count), A(offsets_count));
	return status;
}

RGL_API rgl_status_t rgl_node_rays_transform(rgl_node_t* node, const rgl_mat3x4f* transform)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(transform != nullptr);
		createOrUpdateNode<TransformRaysNode>(node, "TransformRaysNode", false,
		                                      [&](TransformRaysNode& n) { n.setParameters(toMat(transform)); });
	});
	TAPE_RECORD(rgl_node_rays_transform, D(node), R(transform, 1));
	return status;
}

// ============================================================================================ points nodes
RGL_API rgl_status_t rgl_node_points_transform(rgl_node_t* node, const rgl_mat3x4f* transform)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(transform != nullptr);
		createOrUpdateNode<TransformPointsNode>(node, "TransformPointsNode", false,
		                                        [&](TransformPointsNode& n) { n.setParameters(toMat(transform)); });
	});
	TAPE_RECORD(rgl_node_points_transform, D(node), R(transform, 1));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace(rgl_node_t* node, rgl_scene_t scene)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto sc = sceneOf(scene);
		createOrUpdateNode<RaytraceNode>(node, "RaytraceNode", false, [&](RaytraceNode& n) { n.setParameters(sc); });
	});
	TAPE_RECORD(rgl_node_raytrace, D(node), A(scene));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_velocity(rgl_node_t node, const rgl_vec3f* linear_velocity,
                                                          const rgl_vec3f* angular_velocity)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(linear_velocity != nullptr);
		CHECK_ARG(angular_velocity != nullptr);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setVelocity(*reinterpret_cast<const Vec3f*>(linear_velocity), *reinterpret_cast<const Vec3f*>(angular_velocity));
	});
	TAPE_RECORD(rgl_node_raytrace_configure_velocity, A(node), R(linear_velocity, 1), R(angular_velocity, 1));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_distortion(rgl_node_t node, bool enable)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setDistortion(enable);
		n->markDirty();
	});
	TAPE_RECORD(rgl_node_raytrace_configure_distortion, A(node), A(enable));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_non_hits(rgl_node_t node, float nearDistance, float farDistance)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setNonHits(nearDistance, farDistance);
	});
	TAPE_RECORD(rgl_node_raytrace_configure_non_hits, A(node), A(nearDistance), A(farDistance));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_mask(rgl_node_t node, const int8_t* rays_mask, int32_t rays_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(rays_mask != nullptr);
		CHECK_ARG(rays_count > 0);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setMask(rays_mask, rays_count);
		n->markDirty();
	});
	TAPE_RECORD(rgl_node_raytrace_configure_mask, A(node), R(rays_mask, rays_count), A(rays_count));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_beam_divergence(rgl_node_t node, float horizontal_beam_divergence,
                                                                 float vertical_beam_divergence)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG((horizontal_beam_divergence > 0.0f && vertical_beam_divergence > 0.0f) ||
		          (horizontal_beam_divergence == 0.0f && vertical_beam_divergence == 0.0f));
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setBeamDivergence(horizontal_beam_divergence, vertical_beam_divergence);
	});
	TAPE_RECORD(rgl_node_raytrace_configure_beam_divergence, A(node), A(horizontal_beam_divergence), A(vertical_beam_divergence));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_default_intensity(rgl_node_t node, float default_intensity)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(default_intensity >= 0.0f);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setDefaultIntensity(default_intensity);
	});
	TAPE_RECORD(rgl_node_raytrace_configure_default_intensity, A(node), A(default_intensity));
	return status;
}

RGL_API rgl_status_t rgl_node_raytrace_configure_return_mode(rgl_node_t node, rgl_return_mode_t return_mode)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(return_mode != RGL_RETURN_UNKNOWN);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setReturnMode(return_mode);
		if (n->hasCtx()) {
			n->getCtx()->markNodesDirty();
		}
	});
	TAPE_RECORD(rgl_node_raytrace_configure_return_mode, A(node), A(return_mode));
	return status;
}

RGL_API rgl_status_t rgl_node_points_format(rgl_node_t* node, const rgl_field_t* fields, int32_t field_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(fields != nullptr);
		CHECK_ARG(field_count > 0);
		std::vector<rgl_field_t> f(fields, fields + field_count);
		createOrUpdateNode<FormatPointsNode>(node, "FormatPointsNode", true, [&](FormatPointsNode& n) { n.setParameters(f); });
	});
	TAPE_RECORD(rgl_node_points_format, D(node), R(fields, field_count), A(field_count));
	return status;
}

RGL_API rgl_status_t rgl_node_points_yield(rgl_node_t* node, const rgl_field_t* fields, int32_t field_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(fields != nullptr);
		CHECK_ARG(field_count > 0);
		std::vector<rgl_field_t> f(fields, fields + field_count);
		createOrUpdateNode<YieldPointsNode>(node, "YieldPointsNode", true, [&](YieldPointsNode& n) { n.setParameters(f); });
	});
	TAPE_RECORD(rgl_node_points_yield, D(node), R(fields, field_count), A(field_count));
	return status;
}

RGL_API rgl_status_t rgl_node_points_compact_by_field(rgl_node_t* node, rgl_field_t field)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(field == RGL_FIELD_IS_HIT_I32 || field == RGL_FIELD_IS_GROUND_I32);
		createOrUpdateNode<CompactByFieldPointsNode>(node, "CompactByFieldPointsNode", false,
		                                             [&](CompactByFieldPointsNode& n) { n.setParameters(field); });
	});
	TAPE_RECORD(rgl_node_points_compact_by_field, D(node), A(field));
	return status;
}

RGL_API rgl_status_t rgl_node_points_compact(rgl_node_t* node) { return rgl_node_points_compact_by_field(node, RGL_FIELD_IS_HIT_I32); }

RGL_API rgl_status_t rgl_node_points_yield_configure_prefetch(rgl_node_t node, bool enable)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto n = nodeAs<YieldPointsNode>(node, "YieldPointsNode");
		syncNodeGraph(n);
		n->setPrefetch(enable);
	});
}

RGL_API rgl_status_t rgl_node_points_radar_bin(rgl_node_t* node, float range_step, float range_max, float azimuth_step,
                                               float azimuth_min, float azimuth_max, float ray_solid_angle)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		createOrUpdateNode<RadarBinPointsNode>(node, "RadarBinPointsNode", false, [&](RadarBinPointsNode& n) {
			n.setParameters(range_step, range_max, azimuth_step, azimuth_min, azimuth_max, ray_solid_angle);
		});
	});
}

RGL_API rgl_status_t rgl_node_points_spatial_merge(rgl_node_t* node, const rgl_field_t* fields, int32_t field_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(fields != nullptr);
		CHECK_ARG(field_count > 0);
		throw NotImplementedCall("spatial merge is outside the raytrace-and-point-cloud path");
	});
	TAPE_RECORD(rgl_node_points_spatial_merge, D(node), R(fields, field_count), A(field_count));
	return status;
}

RGL_API rgl_status_t rgl_node_points_temporal_merge(rgl_node_t* node, const rgl_field_t* fields, int32_t field_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(fields != nullptr);
		CHECK_ARG(field_count > 0);
		throw NotImplementedCall("temporal merge is outside the raytrace-and-point-cloud path");
	});
	TAPE_RECORD(rgl_node_points_temporal_merge, D(node), R(fields, field_count), A(field_count));
	return status;
}

RGL_API rgl_status_t rgl_node_points_from_array(rgl_node_t* node, const void* points, int32_t points_count,
                                                const rgl_field_t* fields, int32_t field_count)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(points != nullptr);
		CHECK_ARG(points_count > 0);
		CHECK_ARG(fields != nullptr);
		CHECK_ARG(field_count > 0);
		std::vector<rgl_field_t> f(fields, fields + field_count);
		createOrUpdateNode<FromArrayPointsNode>(node, "FromArrayPointsNode", true,
		                                        [&](FromArrayPointsNode& n) { n.setParameters(points, points_count, f); });
	});
	TAPE_RECORD(rgl_node_points_from_array, D(node), R(static_cast<const uint8_t*>(points), static_cast<int64_t>(points_count) * recordBytes(fields, field_count)), A(points_count), R(fields, field_count), A(field_count));
	return status;
}

RGL_API rgl_status_t rgl_node_points_radar_postprocess(rgl_node_t* node, const rgl_radar_scope_t* radar_scopes,
                                                       int32_t radar_scopes_count, float ray_azimuth_step,
                                                       float ray_elevation_step, float frequency, float power_transmitted,
                                                       float cumulative_device_gain, float received_noise_mean,
                                                       float received_noise_st_dev)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(radar_scopes != nullptr);
		CHECK_ARG(radar_scopes_count > 0);
		CHECK_ARG(ray_azimuth_step > 0);
		CHECK_ARG(ray_elevation_step > 0);
		CHECK_ARG(frequency > 0);
		CHECK_ARG(power_transmitted > 0);
		CHECK_ARG(cumulative_device_gain > 0);
		CHECK_ARG(received_noise_st_dev > 0);
		for (int i = 0; i < radar_scopes_count; ++i) {
			CHECK_ARG(radar_scopes[i].begin_distance >= 0);
			CHECK_ARG(radar_scopes[i].end_distance >= 0);
			CHECK_ARG(radar_scopes[i].distance_separation_threshold >= 0);
			CHECK_ARG(radar_scopes[i].radial_speed_separation_threshold >= 0);
			CHECK_ARG(radar_scopes[i].azimuth_separation_threshold >= 0);
			CHECK_ARG(radar_scopes[i].end_distance >= radar_scopes[i].begin_distance);
		}
		std::vector<RadarScope> scopes;
		for (int i = 0; i < radar_scopes_count; ++i) {
			const rgl_radar_scope_t& r = radar_scopes[i];
			scopes.push_back(RadarScope{r.begin_distance, r.end_distance, r.distance_separation_threshold,
			                            r.radial_speed_separation_threshold, r.azimuth_separation_threshold});
		}
		createOrUpdateNode<RadarPostprocessPointsNode>(node, "RadarPostprocessPointsNode", false,
		                                               [&](RadarPostprocessPointsNode& n) {
			                                               n.setParameters(scopes, ray_azimuth_step, ray_elevation_step,
			                                                               frequency, power_transmitted,
			                                                               cumulative_device_gain, received_noise_mean,
			                                                               received_noise_st_dev);
		                                               });
	});
	TAPE_RECORD(rgl_node_points_radar_postprocess, D(node), R(radar_scopes, radar_scopes_count), A(radar_scopes_count), A(ray_azimuth_step), A(ray_elevation_step), A(frequency), A(power_transmitted), A(cumulative_device_gain), A(received_noise_mean), A(received_noise_st_dev));
	return status;
}

RGL_API rgl_status_t rgl_node_points_radar_track_objects(rgl_node_t* node, float object_distance_threshold,
                                                         float object_azimuth_threshold, float object_elevation_threshold,
                                                         float object_radial_speed_threshold, float max_matching_distance,
                                                         float max_prediction_time_frame, float movement_sensitivity)
{
	rgl_status_t status = safeCall([&]() { throw NotImplementedCall("radar object tracking is outside this library"); });
	TAPE_RECORD(rgl_node_points_radar_track_objects, D(node), A(object_distance_threshold), A(object_azimuth_threshold), A(object_elevation_threshold), A(object_radial_speed_threshold), A(max_matching_distance), A(max_prediction_time_frame), A(movement_sensitivity));
	return status;
}

RGL_API rgl_status_t rgl_node_points_radar_set_classes(rgl_node_t node, const int32_t* entity_ids,
                                                       const rgl_radar_object_class_t* object_classes, int32_t count)
{
	rgl_status_t status = safeCall([&]() { throw NotImplementedCall("radar object tracking is outside this library"); });
	TAPE_RECORD(rgl_node_points_radar_set_classes, A(node), R(entity_ids, count), R(object_classes, count), A(count));
	return status;
}

RGL_API rgl_status_t rgl_node_points_filter_ground(rgl_node_t* node, const rgl_vec3f* sensor_up_vector,
                                                   float ground_angle_threshold)
{
	rgl_status_t status = safeCall([&]() { throw NotImplementedCall("ground filtering is outside the raytrace-and-point-cloud path"); });
	TAPE_RECORD(rgl_node_points_filter_ground, D(node), R(sensor_up_vector, 1), A(ground_angle_threshold));
	return status;
}

// ============================================================================================ noise nodes
RGL_API rgl_status_t rgl_node_gaussian_noise_angular_ray(rgl_node_t* node, float mean, float st_dev, rgl_axis_t rotation_axis)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(st_dev >= 0);
		CHECK_ARG((rotation_axis == RGL_AXIS_X) || (rotation_axis == RGL_AXIS_Y) || (rotation_axis == RGL_AXIS_Z));
		createOrUpdateNode<GaussianNoiseAngularRaysNode>(node, "GaussianNoiseAngularRaysNode", false,
		                                                 [&](GaussianNoiseAngularRaysNode& n) {
			                                                 n.setParameters(mean, st_dev, rotation_axis);
		                                                 });
	});
	TAPE_RECORD(rgl_node_gaussian_noise_angular_ray, D(node), A(mean), A(st_dev), A(rotation_axis));
	return status;
}

RGL_API rgl_status_t rgl_node_gaussian_noise_angular_hitpoint(rgl_node_t* node, float mean, float st_dev,
                                                              rgl_axis_t rotation_axis)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(st_dev >= 0);
		CHECK_ARG((rotation_axis == RGL_AXIS_X) || (rotation_axis == RGL_AXIS_Y) || (rotation_axis == RGL_AXIS_Z));
		createOrUpdateNode<GaussianNoiseAngularHitpointNode>(node, "GaussianNoiseAngularHitpointNode", false,
		                                                     [&](GaussianNoiseAngularHitpointNode& n) {
			                                                     n.setParameters(mean, st_dev, rotation_axis);
		                                                     });
	});
	TAPE_RECORD(rgl_node_gaussian_noise_angular_hitpoint, D(node), A(mean), A(st_dev), A(rotation_axis));
	return status;
}

RGL_API rgl_status_t rgl_node_gaussian_noise_distance(rgl_node_t* node, float mean, float st_dev_base,
                                                      float st_dev_rise_per_meter)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(st_dev_base >= 0);
		CHECK_ARG(st_dev_rise_per_meter >= 0);
		createOrUpdateNode<GaussianNoiseDistanceNode>(
		    node, "GaussianNoiseDistanceNode", false,
		    [&](GaussianNoiseDistanceNode& n) { n.setParameters(mean, st_dev_base, st_dev_rise_per_meter); });
	});
	TAPE_RECORD(rgl_node_gaussian_noise_distance, D(node), A(mean), A(st_dev_base), A(st_dev_rise_per_meter));
	return status;
}

// ============================================================================================ graph
RGL_API rgl_status_t rgl_node_is_alive(rgl_node_t node, bool* out_alive)
{
	return safeCall([&]() {
		CHECK_ARG(out_alive != nullptr);
		*out_alive = nodeAlive(node);
	});
}

RGL_API rgl_status_t rgl_graph_run(rgl_node_t raw_node)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(raw_node != nullptr);
		auto node = lookupNode(raw_node);
		if (!node->hasCtx()) {
			GraphRunCtx::createAndAttach(node);
		}
		node->getCtx()->executeAsync();
	});
	TAPE_RECORD(rgl_graph_run, A(raw_node));
	return status;
}

RGL_API rgl_status_t rgl_graph_run_batch(const rgl_node_t* raw_nodes, int32_t node_count)
{
	return safeCall([&]() {
		CHECK_ARG(node_count >= 0);
		CHECK_ARG(raw_nodes != nullptr || node_count == 0);
		std::vector<rgl::Node::Ptr> entries;
		entries.reserve(static_cast<size_t>(node_count));
		for (int32_t i = 0; i < node_count; ++i) {
			CHECK_ARG(raw_nodes[i] != nullptr);
			entries.push_back(lookupNode(raw_nodes[i]));
		}
		GraphBatch::instance().run(entries);
	});
}

RGL_API rgl_status_t rgl_graph_get_result_sizes(const rgl_node_t* raw_nodes, int32_t node_count, rgl_field_t field,
                                                int32_t* out_counts)
{
	return safeCall([&]() {
		CHECK_ARG(node_count >= 0);
		CHECK_ARG((raw_nodes != nullptr && out_counts != nullptr) || node_count == 0);
		for (int32_t i = 0; i < node_count; ++i) {
			CHECK_ARG(raw_nodes[i] != nullptr);
			auto pn = nodeAs<IPointsNode>(raw_nodes[i], "IPointsNode");
			pn->waitForCounts();
			checkOverflow(pn);
			(void) pn->fieldPointSize(field); // an unknown field fails like rgl_graph_get_result_size
			out_counts[i] = pn->hostCount();
		}
	});
}

RGL_API rgl_status_t rgl_graph_batch_set_kernel_timing(bool enable)
{
	return safeCall([&]() { GraphBatch::instance().setTiming(enable); });
}

RGL_API rgl_status_t rgl_graph_batch_get_kernel_timing(double* out_total_ms, int32_t* out_launches)
{
	return safeCall([&]() {
		CHECK_ARG(out_total_ms != nullptr && out_launches != nullptr);
		GraphBatch::instance().getTiming(out_total_ms, out_launches);
	});
}

RGL_API rgl_status_t rgl_graph_batch_configure_capture(bool enable)
{
	return safeCall([&]() { GraphBatch::instance().setGraphs(enable); });
}

RGL_API rgl_status_t rgl_graph_batch_get_capture_stats(int64_t* out_captures, int64_t* out_replays)
{
	return safeCall([&]() {
		CHECK_ARG(out_captures != nullptr && out_replays != nullptr);
		*out_captures = GraphBatch::instance().graphCaptures;
		*out_replays = GraphBatch::instance().graphReplays;
	});
}

RGL_API rgl_status_t rgl_graph_batch_get_fused_stages(int64_t* out_stages)
{
	return safeCall([&]() {
		CHECK_ARG(out_stages != nullptr);
		*out_stages = GraphBatch::instance().fusedStages;
	});
}

RGL_API rgl_status_t rgl_graph_batch_configure_autotune(bool enable)
{
	return safeCall([&]() { GraphBatch::instance().setAutotune(enable); });
}

RGL_API rgl_status_t rgl_graph_batch_force_variants(int32_t raytrace_variant, int32_t compact_variant)
{
	return safeCall([&]() { GraphBatch::instance().forceVariants(raytrace_variant, compact_variant); });
}

RGL_API rgl_status_t rgl_graph_batch_get_variants(int32_t* out_raytrace_variant, int32_t* out_compact_variant)
{
	return safeCall([&]() {
		CHECK_ARG(out_raytrace_variant != nullptr && out_compact_variant != nullptr);
		GraphBatch::instance().autotuneChoice(out_raytrace_variant, out_compact_variant);
	});
}

RGL_API rgl_status_t rgl_graph_destroy(rgl_node_t raw_node)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(raw_node != nullptr);
		auto node = lookupNode(raw_node);
		if (node->hasCtx()) {
			node->getCtx()->detachAndDestroy();
		}
		for (auto& n : node->disconnectConnectedNodes()) {
			releaseNode(n.get());
		}
	});
	TAPE_RECORD(rgl_graph_destroy, A(raw_node));
	return status;
}

RGL_API rgl_status_t rgl_graph_get_result_size(rgl_node_t node, rgl_field_t field, int32_t* out_count, int32_t* out_size_of)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto pn = nodeAs<IPointsNode>(node, "IPointsNode");
		pn->waitForCounts();
		checkOverflow(pn);
		const int32_t count = pn->hostCount();
		const int32_t size = static_cast<int32_t>(pn->fieldPointSize(field));
		if (out_count != nullptr) {
			*out_count = count;
		}
		if (out_size_of != nullptr) {
			*out_size_of = size;
		}
	});
	TAPE_RECORD(rgl_graph_get_result_size, A(node), A(field), V(out_count), V(out_size_of));
	return status;
}

RGL_API rgl_status_t rgl_graph_get_result_data(rgl_node_t node, rgl_field_t field, void* dst)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(dst != nullptr);
		auto pn = nodeAs<IPointsNode>(node, "IPointsNode");
		if (!pn->hasField(field)) {
			throw InvalidPipeline("node " + pn->getName() + " does not provide field " + fieldName(field));
		}
		pn->waitForResults();
		checkOverflow(pn);
		const size_t bytes = static_cast<size_t>(pn->hostCount()) * pn->fieldPointSize(field);
		if (bytes > 0) {
			auto yield = std::dynamic_pointer_cast<YieldPointsNode>(pn);
			const void* prefetched = yield ? yield->hostData(field) : nullptr;
			if (prefetched != nullptr) {
				hostCopy(dst, prefetched, bytes); // written by the run's own copy kernel
			} else {
				deviceToHost(dst, pn->fieldData(field), bytes);
			}
		}
	});
	TAPE_RECORD(rgl_graph_get_result_data, A(node), A(field), A(static_cast<const void*>(dst)));
	return status;
}

RGL_API rgl_status_t rgl_graph_node_add_child(rgl_node_t parent, rgl_node_t child)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(parent != nullptr);
		CHECK_ARG(child != nullptr);
		CHECK_ARG(parent != child);
		lookupNode(parent)->addChild(lookupNode(child));
	});
	TAPE_RECORD(rgl_graph_node_add_child, A(parent), A(child));
	return status;
}

RGL_API rgl_status_t rgl_graph_node_remove_child(rgl_node_t parent, rgl_node_t child)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(parent != nullptr);
		CHECK_ARG(child != nullptr);
		lookupNode(parent)->removeChild(lookupNode(child));
	});
	TAPE_RECORD(rgl_graph_node_remove_child, A(parent), A(child));
	return status;
}

RGL_API rgl_status_t rgl_graph_node_set_priority(rgl_node_t node, int32_t priority)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto n = lookupNode(node);
		syncNodeGraph(n);
		n->setPriority(priority);
	});
	TAPE_RECORD(rgl_graph_node_set_priority, A(node), A(priority));
	return status;
}

RGL_API rgl_status_t rgl_graph_node_get_priority(rgl_node_t node, int32_t* out_priority)
{
	rgl_status_t status = safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(out_priority != nullptr);
		*out_priority = lookupNode(node)->getPriority();
	});
	TAPE_RECORD(rgl_graph_node_get_priority, A(node), V(out_priority));
	return status;
}

// ============================================================================================ extensions
RGL_API rgl_status_t rgl_graph_get_execution_order(rgl_node_t node, rgl_node_t* out_nodes, int32_t* inout_count)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(inout_count != nullptr);
		CHECK_ARG(*inout_count >= 0);
		CHECK_ARG(out_nodes != nullptr || *inout_count == 0);
		auto n = lookupNode(node);
		const auto order = rgl::GraphRunCtx::findExecutionOrder(n->connectedComponent());
		const int32_t capacity = *inout_count;
		for (int32_t i = 0; i < capacity && i < static_cast<int32_t>(order.size()); ++i) {
			out_nodes[i] = handleOf(order[i]);
		}
		*inout_count = static_cast<int32_t>(order.size());
	});
}

RGL_API rgl_status_t rgl_node_gaussian_noise_set_seed(rgl_node_t node, uint64_t seed)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto n = lookupNode(node);
		syncNodeGraph(n);
		NoiseStream* s = nullptr;
		if (auto a = std::dynamic_pointer_cast<GaussianNoiseAngularRaysNode>(n)) {
			s = &a->rng;
		} else if (auto b = std::dynamic_pointer_cast<GaussianNoiseAngularHitpointNode>(n)) {
			s = &b->rng;
		} else if (auto c = std::dynamic_pointer_cast<GaussianNoiseDistanceNode>(n)) {
			s = &c->rng;
		} else if (auto d = std::dynamic_pointer_cast<RadarPostprocessPointsNode>(n)) {
			s = &d->rng;
		}
		if (s == nullptr) {
			throw InvalidAPIObject("RGL API Error: Node type mismatch: expected a Gaussian noise node, got " + n->getName());
		}
		s->seed = seed;
		s->run = 0;
	});
}

RGL_API rgl_status_t rgl_node_gaussian_noise_get_stream(rgl_node_t node, uint64_t* out_seed, uint32_t* out_run,
                                                        uint32_t* out_tag)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(out_seed != nullptr && out_run != nullptr && out_tag != nullptr);
		auto n = lookupNode(node);
		const NoiseStream* s = nullptr;
		if (auto a = std::dynamic_pointer_cast<GaussianNoiseAngularRaysNode>(n)) {
			s = &a->rng;
		} else if (auto b = std::dynamic_pointer_cast<GaussianNoiseAngularHitpointNode>(n)) {
			s = &b->rng;
		} else if (auto c = std::dynamic_pointer_cast<GaussianNoiseDistanceNode>(n)) {
			s = &c->rng;
		} else if (auto d = std::dynamic_pointer_cast<RadarPostprocessPointsNode>(n)) {
			s = &d->rng;
		}
		if (s == nullptr) {
			throw InvalidAPIObject("RGL API Error: Node type mismatch: expected a Gaussian noise node, got " + n->getName());
		}
		*out_seed = s->seed;
		*out_run = s->run;
		*out_tag = s->tag;
	});
}

RGL_API rgl_status_t rgl_scene_set_tlas_refit(rgl_scene_t scene, bool refit)
{
	return safeCall([&]() {
		GraphRunCtx::synchronizeAll();
		sceneOf(scene)->tlasUpdate = refit ? rgl::TlasUpdate::Refit : rgl::TlasUpdate::Rebuild;
	});
}

RGL_API rgl_status_t rgl_scene_create(rgl_scene_t* out_scene)
{
	return safeCall([&]() {
		CHECK_ARG(out_scene != nullptr);
		auto sc = std::make_shared<rgl::Scene>();
		Registry<rgl::Scene>::add(sc);
		*out_scene = reinterpret_cast<rgl_scene_t>(sc.get());
	});
}

RGL_API rgl_status_t rgl_scene_destroy(rgl_scene_t scene)
{
	return safeCall([&]() {
		CHECK_ARG(scene != nullptr);
		GraphRunCtx::synchronizeAll();
		auto sc = Registry<rgl::Scene>::get(scene, "Scene");
		for (auto it = Registry<rgl::Entity>::map().begin(); it != Registry<rgl::Entity>::map().end();) {
			it = it->second->scene == sc.get() ? Registry<rgl::Entity>::map().erase(it) : std::next(it);
		}
		sc->clear();
		Registry<rgl::Scene>::release(scene, "Scene");
	});
}

// ============================================================================================ serving
RGL_API rgl_status_t rgl_server_create(rgl_server_t* out_server, int32_t slots, int32_t max_batch_graphs,
                                       int64_t hbm_budget_bytes, bool batching)
{
	return safeCall([&]() {
		CHECK_ARG(out_server != nullptr);
		CHECK_ARG(slots > 0 && max_batch_graphs > 0);
		auto sv = std::make_shared<rgl::Server>(slots, max_batch_graphs, hbm_budget_bytes, batching);
		Registry<rgl::Server>::add(sv);
		*out_server = reinterpret_cast<rgl_server_t>(sv.get());
	});
}

RGL_API rgl_status_t rgl_server_destroy(rgl_server_t server)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr);
		Registry<rgl::Server>::get(server, "Server")->drain();
		Registry<rgl::Server>::release(server, "Server");
	});
}

RGL_API rgl_status_t rgl_server_submit(rgl_server_t server, rgl_node_t node, int64_t* out_ticket)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr && node != nullptr && out_ticket != nullptr);
		*out_ticket = Registry<rgl::Server>::get(server, "Server")->submit(lookupNode(node));
	});
}

RGL_API rgl_status_t rgl_server_step(rgl_server_t server, int32_t* out_retired)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr);
		const int n = Registry<rgl::Server>::get(server, "Server")->step();
		if (out_retired != nullptr) {
			*out_retired = n;
		}
	});
}

RGL_API rgl_status_t rgl_server_poll(rgl_server_t server, int64_t ticket, int32_t* out_state, double* out_latency_ms)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr && out_state != nullptr);
		*out_state = Registry<rgl::Server>::get(server, "Server")->state(ticket, out_latency_ms);
	});
}

RGL_API rgl_status_t rgl_server_wait(rgl_server_t server, int64_t ticket, double* out_latency_ms)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr);
		Registry<rgl::Server>::get(server, "Server")->wait(ticket, out_latency_ms);
	});
}

RGL_API rgl_status_t rgl_server_get_stats(rgl_server_t server, rgl_server_stats_t* out_stats)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr && out_stats != nullptr);
		const rgl::ServerStats st = Registry<rgl::Server>::get(server, "Server")->stats();
		out_stats->submitted = st.submitted;
		out_stats->completed = st.completed;
		out_stats->batches = st.batches;
		out_stats->rays_completed = st.raysCompleted;
		out_stats->evictions = st.evictions;
		out_stats->deferred = st.deferred;
		out_stats->failed = st.failed;
		out_stats->resident_bytes = st.residentBytes;
		out_stats->p50_ms = st.p50Ms;
		out_stats->p99_ms = st.p99Ms;
		out_stats->mean_ms = st.meanMs;
		out_stats->max_ms = st.maxMs;
	});
}

RGL_API rgl_status_t rgl_server_reset_stats(rgl_server_t server)
{
	return safeCall([&]() {
		CHECK_ARG(server != nullptr);
		Registry<rgl::Server>::get(server, "Server")->resetStats();
	});
}

RGL_API rgl_status_t rgl_configure_bvh_format(int32_t format)
{
	return safeCall([&]() {
		CHECK_ARG(format >= 0 && format < rgl::kBvhFormats);
		rgl::Scene::setBvhFormat(format);
	});
}

RGL_API rgl_status_t rgl_get_bvh_format(int32_t* out_format)
{
	return safeCall([&]() {
		CHECK_ARG(out_format != nullptr);
		*out_format = rgl::Scene::bvhFormat();
	});
}

RGL_API rgl_status_t rgl_node_raytrace_traversal_stats(rgl_node_t node, int32_t* nodes_visited, int32_t* tris_tested,
                                                       int32_t* instances_entered, int32_t ray_count)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(nodes_visited != nullptr && tris_tested != nullptr && instances_entered != nullptr);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		n->waitForResults();
		CHECK_ARG(ray_count == n->capacity() / std::max(1, n->returnCount())); // the kernel writes one entry per ray
		const int rays = ray_count;
		DeviceBuffer a, b, c;
		n->traversalStats(a.ensureT<int32_t>(rays), b.ensureT<int32_t>(rays), c.ensureT<int32_t>(rays));
		RGL_HIP(hipStreamSynchronize(n->stream()));
		RGL_HIP(hipMemcpy(nodes_visited, a.get(), sizeof(int32_t) * rays, hipMemcpyDeviceToHost));
		RGL_HIP(hipMemcpy(tris_tested, b.get(), sizeof(int32_t) * rays, hipMemcpyDeviceToHost));
		RGL_HIP(hipMemcpy(instances_entered, c.get(), sizeof(int32_t) * rays, hipMemcpyDeviceToHost));
	});
}

RGL_API rgl_status_t rgl_graph_traversal_footprint(const rgl_node_t* raytrace_nodes, int32_t count, int64_t* out_bytes)
{
	return safeCall([&]() {
		CHECK_ARG(raytrace_nodes != nullptr);
		CHECK_ARG(count > 0);
		CHECK_ARG(out_bytes != nullptr);
		std::vector<std::shared_ptr<RaytraceNode>> nodes;
		for (int32_t i = 0; i < count; ++i) {
			nodes.push_back(nodeAs<RaytraceNode>(raytrace_nodes[i], "RaytraceNode"));
			nodes.back()->waitForResults();
		}
		// 2^26 keys (512 MiB): room for 2^25 distinct lines (2 GiB of records) at half load.
		constexpr uint32_t kKeys = 1u << 26;
		DeviceBuffer keys, counters;
		RGL_HIP(hipMemset(keys.ensure(sizeof(unsigned long long) * kKeys), 0, sizeof(unsigned long long) * kKeys));
		RGL_HIP(hipMemset(counters.ensure(16), 0, 16));
		auto* full = reinterpret_cast<uint32_t*>(counters.as<char>() + 8);
		for (auto& n : nodes) {
			n->footprint(keys.as<unsigned long long>(), kKeys - 1, full);
			RGL_HIP(hipStreamSynchronize(n->stream()));
		}
		gpu::countNonzeroKeys(nullptr, keys.as<unsigned long long>(), kKeys, counters.as<unsigned long long>());
		RGL_HIP(hipGetLastError());
		unsigned long long host[2] = {0, 0};
		RGL_HIP(hipMemcpy(host, counters.get(), 16, hipMemcpyDeviceToHost));
		if ((host[1] & 0xffffffffull) != 0 || host[0] > kKeys / 2) {
			throw std::runtime_error("traversal footprint exceeds the measurement's line set");
		}
		*out_bytes = static_cast<int64_t>(host[0]) * 64;
	});
}

RGL_API rgl_status_t rgl_node_raytrace_brute_force(rgl_node_t node, float* out_t, int32_t* out_instance,
                                                   int32_t* out_prim, int32_t ray_count)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(out_t != nullptr && out_instance != nullptr && out_prim != nullptr);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		n->waitForResults();
		CHECK_ARG(ray_count == n->capacity() / std::max(1, n->returnCount())); // the kernel writes one entry per ray
		const int rays = ray_count;
		DeviceBuffer a, b, c;
		n->bruteForce(a.ensureT<float>(rays), b.ensureT<int32_t>(rays), c.ensureT<int32_t>(rays));
		RGL_HIP(hipStreamSynchronize(n->stream())); // the node's stream does not synchronize with the null stream
		RGL_HIP(hipMemcpy(out_t, a.get(), sizeof(float) * rays, hipMemcpyDeviceToHost));
		RGL_HIP(hipMemcpy(out_instance, b.get(), sizeof(int32_t) * rays, hipMemcpyDeviceToHost));
		RGL_HIP(hipMemcpy(out_prim, c.get(), sizeof(int32_t) * rays, hipMemcpyDeviceToHost));
	});
}

RGL_API rgl_status_t rgl_node_raytrace_set_kernel_timing(rgl_node_t node, bool enable)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		auto n = nodeAs<RaytraceNode>(node, "RaytraceNode");
		syncNodeGraph(n);
		n->setTiming(enable);
	});
}

RGL_API rgl_status_t rgl_node_raytrace_get_kernel_timing(rgl_node_t node, double* out_total_ms, int32_t* out_launches)
{
	return safeCall([&]() {
		CHECK_ARG(node != nullptr);
		CHECK_ARG(out_total_ms != nullptr && out_launches != nullptr);
		nodeAs<RaytraceNode>(node, "RaytraceNode")->getTiming(out_total_ms, out_launches);
	});
}

RGL_API rgl_status_t rgl_noise_normals(uint64_t seed, uint32_t run, uint32_t tag, float* out, int32_t count)
{
	return safeCall([&]() {
		CHECK_ARG(out != nullptr);
		CHECK_ARG(count > 0);
		DeviceBuffer d;
		gpu::noiseNormals(nullptr, count, d.ensureT<float>(count), seed, run, tag);
		RGL_HIP(hipGetLastError());
		RGL_HIP(hipDeviceSynchronize());
		RGL_HIP(hipMemcpy(out, d.get(), sizeof(float) * count, hipMemcpyDeviceToHost));
	});
}

RGL_API rgl_status_t rgl_noise_normals_no_alloc(uint64_t seed, uint32_t run, uint32_t tag, float* out, int32_t count)
{
	return safeCall([&]() {
		CHECK_ARG(out != nullptr);
		CHECK_ARG(count > 0);
		gpu::noiseNormalsNoAlloc(count, out, seed, run, tag);
	});
}

// ============================================================================================ tape
// src/api/apiCore.cpp:1456-1530
RGL_API rgl_status_t rgl_tape_record_begin(const char* path)
{
	return safeCall([&]() {
		CHECK_ARG(path != nullptr);
		CHECK_ARG(path[0] != '\0');
		tape::begin(path);
	});
}

RGL_API rgl_status_t rgl_tape_record_end()
{
	return safeCall([&]() { tape::end(); });
}

RGL_API rgl_status_t rgl_tape_record_is_active(bool* is_active)
{
	return safeCall([&]() {
		CHECK_ARG(is_active != nullptr);
		*is_active = tape::active() != nullptr;
	});
}

RGL_API rgl_status_t rgl_tape_play(const char* path)
{
	return safeCall([&]() {
		CHECK_ARG(path != nullptr);
		CHECK_ARG(path[0] != '\0');
		if (tape::active() != nullptr) {
			throw tape::TapeError("rgl_tape_play: recording active");
		}
		tape::play(path);
	});
}
