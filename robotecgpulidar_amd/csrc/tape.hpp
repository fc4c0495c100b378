// Tapes: recording of API call streams and their replay (the reference's format, src/tape/TapeRecorder.hpp:27-151,
// src/tape/TapePlayer.cpp:27-150, src/tape/tapeDefinitions.hpp:19-24).
//
// A tape is two files. `<path>.yaml` is a YAML sequence with one single-key map per call,
//     - rgl_mesh_create: {t: 1234567, a: [94811234, 0, 8, 96, 12]}
// (t: nanoseconds since recording began; a: the arguments). `<path>.bin` holds the array arguments, each padded to
// 16 bytes; an array argument is recorded as its byte offset in that file. Object handles are recorded as their
// addresses and mapped to the live objects of the replay; out-parameters of type int32_t* are recorded by value;
// enums as integers. The YAML is written and read by this file's own emitter and parser for exactly this subset
// (block sequence of `name: {t: int, a: [scalars]}`; scalars plain, "double-quoted", 'single-quoted' or ~).
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include <rgl/api/core.h>

namespace rgl::tape {

// Version of the tape format this library writes and the oldest it reads (src/tape/tapeDefinitions.hpp:19-21).
constexpr int kFormatMajor = 0, kFormatMinor = 16, kFormatPatch = 3;

struct TapeError : std::runtime_error
{
	using std::runtime_error::runtime_error;
};
struct InvalidFilePath : std::invalid_argument
{
	using std::invalid_argument::invalid_argument;
};

// An array argument: `bytes` bytes at `data` go to the binary file.
struct Bytes
{
	const void* data;
	size_t bytes;
};
template<typename T>
Bytes array(const T* p, int64_t count)
{
	return Bytes{p, (p != nullptr && count > 0) ? sizeof(T) * static_cast<size_t>(count) : 0};
}

class Recorder
{
public:
	explicit Recorder(const std::string& path);
	~Recorder();
	// Appends one call; `args` are YAML scalars already formatted (see the scalar helpers below).
	void call(const char* fn, const std::vector<std::string>& args);
	// Writes `b` to the binary file and returns its offset as a scalar.
	std::string bytes(const Bytes& b);

private:
	FILE* yaml = nullptr;
	FILE* bin = nullptr;
	uint64_t binOffset = 0;
	int64_t beginNs = 0;
};

// The active recorder, or nullptr.
Recorder* active();
void begin(const std::string& path);
void end();

// YAML scalars.
std::string scalar(int64_t v);
std::string scalarU(uint64_t v);
std::string scalar(double v, bool single);
std::string scalar(bool v);
std::string scalar(const char* s);
std::string handle(const void* p);

// Replays a tape: every call in order, each at (approximately) its recorded time after the start.
void play(const std::string& path);

} // namespace rgl::tape
