// Host runtime primitives: HIP error checks, growth-only device / pinned buffers, streams, events, exceptions,
// and the point-field registry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <rgl/api/core.h>

#include "math.hpp"

namespace rgl {

// Exception kinds mapped to status codes by the API layer (src/api/apiCommon.hpp:70-105).
struct InvalidAPIObject : std::invalid_argument
{
	using std::invalid_argument::invalid_argument;
};
struct InvalidPipeline : std::invalid_argument
{
	using std::invalid_argument::invalid_argument;
};
struct NotImplemented : std::runtime_error
{
	using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error
{
	using std::runtime_error::runtime_error;
};

void throwHip(hipError_t e, const char* expr, const char* file, int line);
#define RGL_HIP(expr)                                                                                                  \
	do {                                                                                                               \
		hipError_t _e = (expr);                                                                                        \
		if (_e != hipSuccess) {                                                                                        \
			::rgl::throwHip(_e, #expr, __FILE__, __LINE__);                                                            \
		}                                                                                                              \
	} while (0)

// Growth-only device allocation; contents are not preserved across growth.
class DeviceBuffer
{
public:
	DeviceBuffer() = default;
	DeviceBuffer(const DeviceBuffer&) = delete;
	DeviceBuffer& operator=(const DeviceBuffer&) = delete;
	~DeviceBuffer() { release(); }

	void* ensure(size_t bytes)
	{
		if (bytes > cap) {
			release();
			size_t want = bytes < 256 ? 256 : bytes;
			RGL_HIP(hipMalloc(&ptr, want));
			cap = want;
		}
		return ptr;
	}
	template<typename T>
	T* ensureT(size_t count)
	{
		return static_cast<T*>(ensure(count * sizeof(T)));
	}
	void* get() const { return ptr; }
	template<typename T>
	T* as() const
	{
		return static_cast<T*>(ptr);
	}
	size_t capacity() const { return cap; }
	void release()
	{
		if (ptr != nullptr) {
			(void) hipFree(ptr);
			ptr = nullptr;
			cap = 0;
		}
	}

private:
	void* ptr = nullptr;
	size_t cap = 0;
};

class PinnedBuffer
{
public:
	PinnedBuffer() = default;
	PinnedBuffer(const PinnedBuffer&) = delete;
	PinnedBuffer& operator=(const PinnedBuffer&) = delete;
	~PinnedBuffer() { release(); }
	void* ensure(size_t bytes)
	{
		if (bytes > cap) {
			release();
			size_t want = bytes < 256 ? 256 : bytes;
			RGL_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
			cap = want;
		}
		return ptr;
	}
	void* get() const { return ptr; }
	size_t capacity() const { return cap; }
	void release()
	{
		if (ptr != nullptr) {
			(void) hipHostFree(ptr);
			ptr = nullptr;
			cap = 0;
		}
	}

private:
	void* ptr = nullptr;
	size_t cap = 0;
};

// Pinned host memory mapped into the device's address space: kernels write results into it over the bus, and the
// host reads them once the writing work has completed.
class MappedHostBuffer
{
public:
	MappedHostBuffer() = default;
	MappedHostBuffer(const MappedHostBuffer&) = delete;
	MappedHostBuffer& operator=(const MappedHostBuffer&) = delete;
	~MappedHostBuffer() { release(); }
	void* ensure(size_t bytes)
	{
		if (bytes > cap) {
			release();
			const size_t want = bytes < 256 ? 256 : bytes;
			RGL_HIP(hipHostMalloc(&ptr, want, hipHostMallocMapped | hipHostMallocCoherent));
			RGL_HIP(hipHostGetDevicePointer(&dev, ptr, 0));
			cap = want;
		}
		return ptr;
	}
	void* host() const { return ptr; }
	void* device() const { return dev; }
	size_t capacity() const { return cap; }
	void release()
	{
		if (ptr != nullptr) {
			(void) hipHostFree(ptr);
			ptr = dev = nullptr;
			cap = 0;
		}
	}

private:
	void* ptr = nullptr;
	void* dev = nullptr;
	size_t cap = 0;
};

class Stream
{
public:
	Stream() { RGL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
	~Stream() { (void) hipStreamDestroy(s); }
	Stream(const Stream&) = delete;
	Stream& operator=(const Stream&) = delete;
	hipStream_t get() const { return s; }
	void sync() { RGL_HIP(hipStreamSynchronize(s)); }

private:
	hipStream_t s = nullptr;
};

class Event
{
public:
	Event() { RGL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
	~Event() { (void) hipEventDestroy(e); }
	Event(const Event&) = delete;
	Event& operator=(const Event&) = delete;
	hipEvent_t get() const { return e; }
	void record(hipStream_t s)
	{
		RGL_HIP(hipEventRecord(e, s));
		recorded = true;
		reached = false;
	}
	// Waits for the last recorded work; once it has been seen complete, later waits return at once (result queries
	// wait on the same events many times per frame).
	void sync()
	{
		if (recorded && !reached) {
			RGL_HIP(hipEventSynchronize(e));
			reached = true;
		}
	}
	bool wasRecorded() const { return recorded; }

private:
	hipEvent_t e = nullptr;
	bool recorded = false;
	bool reached = false;
};

// ------------------------------------------------------------------------------------------------ fields
// Field -> element size (src/RGLFields.hpp:108-191). 0 = unknown field.
inline size_t fieldSize(rgl_field_t f)
{
	switch (f) {
		case RGL_FIELD_XYZ_VEC3_F32: return 12;
		case RGL_FIELD_INTENSITY_F32: return 4;
		case RGL_FIELD_INTENSITY_U8: return 1;
		case RGL_FIELD_IS_HIT_I32: return 4;
		case RGL_FIELD_IS_GROUND_I32: return 4;
		case RGL_FIELD_RAY_IDX_U32: return 4;
		case RGL_FIELD_ENTITY_ID_I32: return 4;
		case RGL_FIELD_DISTANCE_F32: return 4;
		case RGL_FIELD_AZIMUTH_F32: return 4;
		case RGL_FIELD_ELEVATION_F32: return 4;
		case RGL_FIELD_RING_ID_U16: return 2;
		case RGL_FIELD_RETURN_TYPE_U8: return 1;
		case RGL_FIELD_TIME_STAMP_F64: return 8;
		case RGL_FIELD_TIME_STAMP_U32: return 4;
		case RGL_FIELD_ABSOLUTE_VELOCITY_VEC3_F32: return 12;
		case RGL_FIELD_RELATIVE_VELOCITY_VEC3_F32: return 12;
		case RGL_FIELD_RADIAL_SPEED_F32: return 4;
		case RGL_FIELD_POWER_F32: return 4;
		case RGL_FIELD_RCS_F32: return 4;
		case RGL_FIELD_NOISE_F32: return 4;
		case RGL_FIELD_SNR_F32: return 4;
		case RGL_FIELD_NORMAL_VEC3_F32: return 12;
		case RGL_FIELD_INCIDENT_ANGLE_F32: return 4;
		case RGL_FIELD_RAY_POSE_MAT3x4_F32: return 48;
		case RGL_FIELD_LASER_RETRO_F32: return 4;
		case RGL_FIELD_PADDING_8: return 1;
		case RGL_FIELD_PADDING_16: return 2;
		case RGL_FIELD_PADDING_32: return 4;
		default: return 0;
	}
}

inline size_t fieldSizeChecked(rgl_field_t f)
{
	size_t s = fieldSize(f);
	if (s == 0) {
		throw std::invalid_argument("unknown RGL field " + std::to_string(static_cast<int>(f)));
	}
	return s;
}

inline bool isPadding(rgl_field_t f)
{
	return f == RGL_FIELD_PADDING_8 || f == RGL_FIELD_PADDING_16 || f == RGL_FIELD_PADDING_32;
}

inline size_t pointSize(const std::vector<rgl_field_t>& fields)
{
	size_t s = 0;
	for (auto f : fields) {
		s += fieldSizeChecked(f);
	}
	return s;
}

std::string fieldName(rgl_field_t f);

} // namespace rgl
