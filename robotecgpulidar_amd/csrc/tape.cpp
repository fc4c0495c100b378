#include "tape.hpp"

#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <thread>
#include <unordered_map>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <rgl/api/extensions/tape.h>

#include "graph.hpp"

namespace rgl::tape {

namespace {

int64_t nowNs()
{
	return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
	    .count();
}

std::unique_ptr<Recorder>& recorderSlot()
{
	static std::unique_ptr<Recorder> r;
	return r;
}

} // namespace

// ------------------------------------------------------------------------------------------------ scalars
std::string scalar(int64_t v) { return std::to_string(v); }
std::string scalarU(uint64_t v) { return std::to_string(v); }
std::string scalar(bool v) { return v ? "true" : "false"; }
std::string handle(const void* p) { return std::to_string(reinterpret_cast<uintptr_t>(p)); }

std::string scalar(double v, bool single)
{
	if (std::isnan(v)) {
		return ".nan";
	}
	if (std::isinf(v)) {
		return v > 0 ? ".inf" : "-.inf";
	}
	char buf[64];
	// Enough significant digits to read back the same float / double.
	std::snprintf(buf, sizeof(buf), single ? "%.9g" : "%.17g", v);
	return buf;
}

std::string scalar(const char* s)
{
	if (s == nullptr) {
		return "~";
	}
	std::string out = "\"";
	for (const char* c = s; *c != '\0'; ++c) {
		switch (*c) {
			case '"': out += "\\\""; break;
			case '\\': out += "\\\\"; break;
			case '\n': out += "\\n"; break;
			case '\t': out += "\\t"; break;
			default: out += *c;
		}
	}
	return out + "\"";
}

// ------------------------------------------------------------------------------------------------ recorder
Recorder::Recorder(const std::string& path)
{
	const std::string yamlPath = path + ".yaml", binPath = path + ".bin";
	bin = std::fopen(binPath.c_str(), "wb");
	if (bin == nullptr) {
		throw InvalidFilePath("rgl_tape_record_begin: could not open binary file '" + binPath +
		                      "' due to the error: " + std::strerror(errno));
	}
	yaml = std::fopen(yamlPath.c_str(), "w");
	if (yaml == nullptr) {
		std::fclose(bin);
		throw InvalidFilePath("rgl_tape_record_begin: could not open yaml file '" + yamlPath +
		                      "' due to the error: " + std::strerror(errno));
	}
	beginNs = nowNs();
	// The first record is the library version (src/tape/TapeRecorder.cpp:58-63).
	call("rgl_get_version_info", {scalar(int64_t{RGL_VERSION_MAJOR}), scalar(int64_t{RGL_VERSION_MINOR}),
	                              scalar(int64_t{RGL_VERSION_PATCH})});
}

Recorder::~Recorder()
{
	if (yaml != nullptr) {
		std::fclose(yaml);
	}
	if (bin != nullptr) {
		std::fclose(bin);
	}
}

void Recorder::call(const char* fn, const std::vector<std::string>& args)
{
	std::string line = std::string("- ") + fn + ": {t: " + std::to_string(nowNs() - beginNs) + ", a: [";
	for (size_t i = 0; i < args.size(); ++i) {
		line += (i ? ", " : "") + args[i];
	}
	line += "]}\n";
	if (std::fputs(line.c_str(), yaml) < 0) {
		throw TapeError("Failed to write data to yaml file");
	}
}

std::string Recorder::bytes(const Bytes& b)
{
	const uint64_t at = binOffset;
	if (b.bytes > 0 && std::fwrite(b.data, 1, b.bytes, bin) != b.bytes) {
		throw TapeError("Failed to write data to binary file");
	}
	const size_t pad = (16 - b.bytes % 16) % 16;
	static const char zeros[16] = {};
	if (pad > 0 && std::fwrite(zeros, 1, pad, bin) != pad) {
		throw TapeError("Failed to write data to binary file");
	}
	binOffset += b.bytes + pad;
	return scalarU(at);
}

Recorder* active() { return recorderSlot().get(); }

void begin(const std::string& path)
{
	if (recorderSlot()) {
		throw TapeError("rgl_tape_record_begin: recording already active");
	}
	recorderSlot() = std::make_unique<Recorder>(path);
}

void end()
{
	if (!recorderSlot()) {
		throw TapeError("rgl_tape_record_end: no recording active");
	}
	recorderSlot().reset();
}

// ------------------------------------------------------------------------------------------------ parser
namespace {

struct Call
{
	std::string fn;
	int64_t t = 0;
	std::vector<std::string> args; // scalars, unquoted
};

class Parser
{
public:
	explicit Parser(std::string text) : s(std::move(text)) {}

	std::vector<Call> parse()
	{
		std::vector<Call> calls;
		skipSpaceAndComments();
		if (pos >= s.size()) {
			throw TapeError("Invalid Tape: Empty YAML file detected");
		}
		if (s[pos] != '-' || (pos + 1 < s.size() && s[pos + 1] != ' ' && s[pos + 1] != '\n')) {
			throw TapeError("Unsupported Tape format: Detected outdated format");
		}
		while (pos < s.size()) {
			expect('-');
			Call c;
			c.fn = plainToken(":");
			expect(':');
			expect('{');
			while (true) {
				const std::string key = plainToken(":");
				expect(':');
				if (key == "t") {
					c.t = std::stoll(plainToken(",}"));
				} else if (key == "a") {
					expect('[');
					skipSpace();
					if (peek() != ']') {
						while (true) {
							c.args.push_back(scalarToken());
							skipSpace();
							if (peek() == ',') {
								++pos;
								continue;
							}
							break;
						}
					}
					expect(']');
				} else {
					throw TapeError("Invalid Tape: unknown key '" + key + "' in call " + c.fn);
				}
				skipSpace();
				if (peek() == ',') {
					++pos;
					continue;
				}
				break;
			}
			expect('}');
			calls.push_back(std::move(c));
			skipSpaceAndComments();
		}
		return calls;
	}

private:
	char peek() const { return pos < s.size() ? s[pos] : '\0'; }
	void skipSpace()
	{
		while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n' || s[pos] == '\r')) {
			++pos;
		}
	}
	void skipSpaceAndComments()
	{
		while (true) {
			skipSpace();
			if (peek() == '#') {
				while (pos < s.size() && s[pos] != '\n') {
					++pos;
				}
				continue;
			}
			if (s.compare(pos, 3, "---") == 0) {
				pos += 3;
				continue;
			}
			break;
		}
	}
	void expect(char c)
	{
		skipSpace();
		if (peek() != c) {
			throw TapeError(std::string("Invalid Tape: expected '") + c + "' at byte " + std::to_string(pos));
		}
		++pos;
	}
	// A plain token up to (not including) any of `stops`, trimmed.
	std::string plainToken(const char* stops)
	{
		skipSpace();
		const size_t b = pos;
		while (pos < s.size() && std::strchr(stops, s[pos]) == nullptr && s[pos] != '\n') {
			++pos;
		}
		size_t e = pos;
		while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) {
			--e;
		}
		return s.substr(b, e - b);
	}
	std::string scalarToken()
	{
		skipSpace();
		if (peek() == '"') {
			++pos;
			std::string out;
			while (pos < s.size() && s[pos] != '"') {
				if (s[pos] == '\\' && pos + 1 < s.size()) {
					const char e = s[++pos];
					out += e == 'n' ? '\n' : (e == 't' ? '\t' : e);
				} else {
					out += s[pos];
				}
				++pos;
			}
			++pos;
			return out;
		}
		if (peek() == '\'') {
			++pos;
			std::string out;
			while (pos < s.size()) {
				if (s[pos] == '\'') {
					if (pos + 1 < s.size() && s[pos + 1] == '\'') {
						out += '\'';
						pos += 2;
						continue;
					}
					break;
				}
				out += s[pos++];
			}
			++pos;
			return out;
		}
		return plainToken(",]");
	}

	std::string s;
	size_t pos = 0;
};

int64_t asInt(const std::string& v)
{
	if (v == "true") {
		return 1;
	}
	if (v == "false") {
		return 0;
	}
	size_t used = 0;
	const long long x = std::stoll(v, &used, 0);
	if (used != v.size()) {
		throw TapeError("Invalid Tape: '" + v + "' is not an integer");
	}
	return x;
}
uint64_t asUint(const std::string& v)
{
	size_t used = 0;
	const unsigned long long x = std::stoull(v, &used, 0);
	if (used != v.size()) {
		throw TapeError("Invalid Tape: '" + v + "' is not an integer");
	}
	return x;
}
float asFloat(const std::string& v)
{
	if (v == ".nan" || v == ".NaN" || v == ".NAN") {
		return NAN;
	}
	if (v == ".inf" || v == ".Inf" || v == "+.inf") {
		return INFINITY;
	}
	if (v == "-.inf" || v == "-.Inf") {
		return -INFINITY;
	}
	return std::stof(v);
}
bool asBool(const std::string& v) { return v == "true" || v == "True" || v == "1"; }

// The replay's objects, by recorded handle (src/tape/PlaybackState.hpp:21-61).
struct State
{
	explicit State(const std::string& binPath)
	{
		const int fd = open(binPath.c_str(), O_RDONLY);
		if (fd < 0) {
			throw InvalidFilePath("TAPE: could not open binary file: '" + binPath + "'  due to the error: " +
			                      std::strerror(errno));
		}
		struct stat st
		{};
		if (fstat(fd, &st) < 0) {
			close(fd);
			throw InvalidFilePath("TAPE: couldn't read binary file length");
		}
		size = static_cast<size_t>(st.st_size);
		if (size > 0) {
			void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
			if (m == MAP_FAILED) {
				close(fd);
				throw InvalidFilePath("TAPE: could not mmap binary file: " + binPath);
			}
			data = static_cast<const uint8_t*>(m);
		}
		close(fd);
	}
	~State()
	{
		if (data != nullptr) {
			munmap(const_cast<uint8_t*>(data), size);
		}
	}
	template<typename T>
	const T* ptr(const std::string& offsetScalar, size_t count = 1) const
	{
		const uint64_t off = asUint(offsetScalar);
		const size_t need = (std::is_void_v<T> ? 1 : sizeof(std::conditional_t<std::is_void_v<T>, char, T>)) * count;
		if (data == nullptr) {
			throw TapeError("Trying to get tape binary data but it is empty");
		}
		if (off + need > size) {
			throw TapeError("Tape binary offset with size of requested type (" + std::to_string(off) + "+" +
			                std::to_string(need) + ") out of range (" + std::to_string(size) + ")");
		}
		return reinterpret_cast<const T*>(data + off);
	}
	void clear()
	{
		meshes.clear();
		entities.clear();
		textures.clear();
		nodes.clear();
	}
	const uint8_t* data = nullptr;
	size_t size = 0;
	std::unordered_map<uint64_t, rgl_mesh_t> meshes;
	std::unordered_map<uint64_t, rgl_entity_t> entities;
	std::unordered_map<uint64_t, rgl_texture_t> textures;
	std::unordered_map<uint64_t, rgl_node_t> nodes;
};

using Args = std::vector<std::string>;
using Player = std::function<void(const Args&, State&)>;

void need(const Args& a, size_t n, const char* fn)
{
	if (a.size() < n) {
		throw TapeError(std::string("Invalid Tape: ") + fn + " needs " + std::to_string(n) + " arguments");
	}
}

template<typename Map>
auto lookup(const Map& m, const std::string& key, const char* kind)
{
	auto it = m.find(asUint(key));
	if (it == m.end()) {
		throw TapeError(std::string("Invalid Tape: unknown ") + kind + " " + key);
	}
	return it->second;
}

rgl_node_t nodeOrNull(State& st, const std::string& key)
{
	auto it = st.nodes.find(asUint(key));
	return it == st.nodes.end() ? nullptr : it->second;
}

// Node constructors are create-or-update: the recorded handle maps to the node made by the first call.
template<typename Fn>
Player nodeCall(Fn fn)
{
	return [fn](const Args& a, State& st) {
		need(a, 1, "node call");
		rgl_node_t node = nodeOrNull(st, a[0]);
		fn(&node, a, st);
		if (node != nullptr) {
			st.nodes[asUint(a[0])] = node;
		}
	};
}

const std::map<std::string, Player>& players()
{
	static const std::map<std::string, Player> table = {
	    {"rgl_get_version_info", [](const Args&, State&) {
		     int32_t a, b, c;
		     rgl_get_version_info(&a, &b, &c);
	     }},
	    {"rgl_get_extension_info", [](const Args& a, State&) {
		     need(a, 1, "rgl_get_extension_info");
		     int32_t avail;
		     rgl_get_extension_info(static_cast<rgl_extension_t>(asInt(a[0])), &avail);
	     }},
	    {"rgl_configure_logging", [](const Args& a, State&) {
		     need(a, 3, "rgl_configure_logging");
		     rgl_configure_logging(static_cast<rgl_log_level_t>(asInt(a[0])), a[1] == "~" ? nullptr : a[1].c_str(),
		                           asBool(a[2]));
	     }},
	    {"rgl_cleanup", [](const Args&, State& st) {
		     rgl_cleanup();
		     st.clear();
	     }},
	    {"rgl_mesh_create", [](const Args& a, State& st) {
		     need(a, 5, "rgl_mesh_create");
		     rgl_mesh_t mesh = nullptr;
		     const int32_t nv = static_cast<int32_t>(asInt(a[2])), ni = static_cast<int32_t>(asInt(a[4]));
		     rgl_mesh_create(&mesh, st.ptr<rgl_vec3f>(a[1], nv), nv, st.ptr<rgl_vec3i>(a[3], ni), ni);
		     st.meshes[asUint(a[0])] = mesh;
	     }},
	    {"rgl_mesh_set_texture_coords", [](const Args& a, State& st) {
		     need(a, 3, "rgl_mesh_set_texture_coords");
		     const int32_t n = static_cast<int32_t>(asInt(a[2]));
		     rgl_mesh_set_texture_coords(lookup(st.meshes, a[0], "mesh"), st.ptr<rgl_vec2f>(a[1], n), n);
	     }},
	    {"rgl_mesh_set_bone_weights", [](const Args& a, State& st) {
		     need(a, 3, "rgl_mesh_set_bone_weights");
		     const int32_t n = static_cast<int32_t>(asInt(a[2]));
		     rgl_mesh_set_bone_weights(lookup(st.meshes, a[0], "mesh"), st.ptr<rgl_bone_weights_t>(a[1], n), n);
	     }},
	    {"rgl_mesh_set_restposes", [](const Args& a, State& st) {
		     need(a, 3, "rgl_mesh_set_restposes");
		     const int32_t n = static_cast<int32_t>(asInt(a[2]));
		     rgl_mesh_set_restposes(lookup(st.meshes, a[0], "mesh"), st.ptr<rgl_mat3x4f>(a[1], n), n);
	     }},
	    {"rgl_mesh_destroy", [](const Args& a, State& st) {
		     need(a, 1, "rgl_mesh_destroy");
		     rgl_mesh_destroy(lookup(st.meshes, a[0], "mesh"));
		     st.meshes.erase(asUint(a[0]));
	     }},
	    {"rgl_entity_create", [](const Args& a, State& st) {
		     need(a, 3, "rgl_entity_create");
		     rgl_entity_t e = nullptr;
		     rgl_entity_create(&e, nullptr, lookup(st.meshes, a[2], "mesh"));
		     st.entities[asUint(a[0])] = e;
	     }},
	    {"rgl_entity_destroy", [](const Args& a, State& st) {
		     need(a, 1, "rgl_entity_destroy");
		     rgl_entity_destroy(lookup(st.entities, a[0], "entity"));
		     st.entities.erase(asUint(a[0]));
	     }},
	    {"rgl_entity_set_transform", [](const Args& a, State& st) {
		     need(a, 2, "rgl_entity_set_transform");
		     rgl_entity_set_transform(lookup(st.entities, a[0], "entity"), st.ptr<rgl_mat3x4f>(a[1]));
	     }},
	    {"rgl_entity_set_pose_world", [](const Args& a, State& st) {
		     need(a, 3, "rgl_entity_set_pose_world");
		     const int32_t n = static_cast<int32_t>(asInt(a[2]));
		     rgl_entity_set_pose_world(lookup(st.entities, a[0], "entity"), st.ptr<rgl_mat3x4f>(a[1], n), n);
	     }},
	    {"rgl_entity_set_id", [](const Args& a, State& st) {
		     need(a, 2, "rgl_entity_set_id");
		     rgl_entity_set_id(lookup(st.entities, a[0], "entity"), static_cast<int32_t>(asInt(a[1])));
	     }},
	    {"rgl_entity_set_intensity_texture", [](const Args& a, State& st) {
		     need(a, 2, "rgl_entity_set_intensity_texture");
		     rgl_entity_set_intensity_texture(lookup(st.entities, a[0], "entity"), lookup(st.textures, a[1], "texture"));
	     }},
	    {"rgl_entity_set_laser_retro", [](const Args& a, State& st) {
		     need(a, 2, "rgl_entity_set_laser_retro");
		     rgl_entity_set_laser_retro(lookup(st.entities, a[0], "entity"), asFloat(a[1]));
	     }},
	    {"rgl_entity_apply_external_animation", [](const Args& a, State& st) {
		     need(a, 3, "rgl_entity_apply_external_animation");
		     const int32_t n = static_cast<int32_t>(asInt(a[2]));
		     rgl_entity_apply_external_animation(lookup(st.entities, a[0], "entity"), st.ptr<rgl_vec3f>(a[1], n), n);
	     }},
	    {"rgl_texture_create", [](const Args& a, State& st) {
		     need(a, 4, "rgl_texture_create");
		     rgl_texture_t t = nullptr;
		     const int32_t w = static_cast<int32_t>(asInt(a[2])), h = static_cast<int32_t>(asInt(a[3]));
		     rgl_texture_create(&t, st.ptr<uint8_t>(a[1], static_cast<size_t>(std::max(w, 0)) * std::max(h, 0)), w, h);
		     st.textures[asUint(a[0])] = t;
	     }},
	    {"rgl_texture_destroy", [](const Args& a, State& st) {
		     need(a, 1, "rgl_texture_destroy");
		     rgl_texture_destroy(lookup(st.textures, a[0], "texture"));
		     st.textures.erase(asUint(a[0]));
	     }},
	    {"rgl_scene_set_time", [](const Args& a, State&) {
		     need(a, 2, "rgl_scene_set_time");
		     rgl_scene_set_time(nullptr, asUint(a[1]));
	     }},
	    {"rgl_graph_run", [](const Args& a, State& st) {
		     need(a, 1, "rgl_graph_run");
		     rgl_graph_run(lookup(st.nodes, a[0], "node"));
	     }},
	    {"rgl_graph_destroy", [](const Args& a, State& st) {
		     need(a, 1, "rgl_graph_destroy");
		     rgl_node_t n = lookup(st.nodes, a[0], "node");
		     // Every node of the destroyed graph leaves the map (src/api/apiCore.cpp:569-600).
		     std::vector<uint64_t> gone;
		     if (nodeAlive(n)) {
			     std::set<const void*> graph;
			     for (auto& g : lookupNode(n)->connectedComponent()) {
				     graph.insert(static_cast<const void*>(g.get()));
			     }
			     for (auto& [key, live] : st.nodes) {
				     if (graph.count(static_cast<const void*>(live))) {
					     gone.push_back(key);
				     }
			     }
		     }
		     rgl_graph_destroy(n);
		     for (uint64_t k : gone) {
			     st.nodes.erase(k);
		     }
	     }},
	    {"rgl_graph_get_result_size", [](const Args& a, State& st) {
		     need(a, 2, "rgl_graph_get_result_size");
		     int32_t count = 0, size = 0;
		     rgl_graph_get_result_size(lookup(st.nodes, a[0], "node"), static_cast<rgl_field_t>(asInt(a[1])), &count,
		                               &size);
	     }},
	    {"rgl_graph_get_result_data", [](const Args& a, State& st) {
		     need(a, 2, "rgl_graph_get_result_data");
		     rgl_node_t n = lookup(st.nodes, a[0], "node");
		     const auto f = static_cast<rgl_field_t>(asInt(a[1]));
		     int32_t count = 0, size = 0;
		     if (rgl_graph_get_result_size(n, f, &count, &size) == RGL_SUCCESS && count > 0) {
			     std::vector<char> tmp(static_cast<size_t>(count) * size);
			     rgl_graph_get_result_data(n, f, tmp.data());
		     }
	     }},
	    {"rgl_graph_node_add_child", [](const Args& a, State& st) {
		     need(a, 2, "rgl_graph_node_add_child");
		     rgl_graph_node_add_child(lookup(st.nodes, a[0], "node"), lookup(st.nodes, a[1], "node"));
	     }},
	    {"rgl_graph_node_remove_child", [](const Args& a, State& st) {
		     need(a, 2, "rgl_graph_node_remove_child");
		     rgl_graph_node_remove_child(lookup(st.nodes, a[0], "node"), lookup(st.nodes, a[1], "node"));
	     }},
	    {"rgl_graph_node_set_priority", [](const Args& a, State& st) {
		     need(a, 2, "rgl_graph_node_set_priority");
		     rgl_graph_node_set_priority(lookup(st.nodes, a[0], "node"), static_cast<int32_t>(asInt(a[1])));
	     }},
	    {"rgl_graph_node_get_priority", [](const Args& a, State& st) {
		     need(a, 1, "rgl_graph_node_get_priority");
		     int32_t p = 0;
		     rgl_graph_node_get_priority(lookup(st.nodes, a[0], "node"), &p);
	     }},
	    {"rgl_node_rays_from_mat3x4f", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_rays_from_mat3x4f");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_rays_from_mat3x4f(n, st.ptr<rgl_mat3x4f>(a[1], c), c);
	     })},
	    {"rgl_node_rays_set_ring_ids", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_rays_set_ring_ids");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_rays_set_ring_ids(n, st.ptr<int32_t>(a[1], c), c);
	     })},
	    {"rgl_node_rays_set_range", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_rays_set_range");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_rays_set_range(n, st.ptr<rgl_vec2f>(a[1], c), c);
	     })},
	    {"rgl_node_rays_set_time_offsets", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_rays_set_time_offsets");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_rays_set_time_offsets(n, st.ptr<float>(a[1], c), c);
	     })},
	    {"rgl_node_rays_transform", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 2, "rgl_node_rays_transform");
		     rgl_node_rays_transform(n, st.ptr<rgl_mat3x4f>(a[1]));
	     })},
	    {"rgl_node_points_transform", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 2, "rgl_node_points_transform");
		     rgl_node_points_transform(n, st.ptr<rgl_mat3x4f>(a[1]));
	     })},
	    {"rgl_node_raytrace", nodeCall([](rgl_node_t* n, const Args&, State&) { rgl_node_raytrace(n, nullptr); })},
	    {"rgl_node_raytrace_configure_velocity", [](const Args& a, State& st) {
		     need(a, 3, "rgl_node_raytrace_configure_velocity");
		     rgl_node_raytrace_configure_velocity(lookup(st.nodes, a[0], "node"), st.ptr<rgl_vec3f>(a[1]),
		                                          st.ptr<rgl_vec3f>(a[2]));
	     }},
	    {"rgl_node_raytrace_configure_distortion", [](const Args& a, State& st) {
		     need(a, 2, "rgl_node_raytrace_configure_distortion");
		     rgl_node_raytrace_configure_distortion(lookup(st.nodes, a[0], "node"), asBool(a[1]));
	     }},
	    {"rgl_node_raytrace_configure_non_hits", [](const Args& a, State& st) {
		     need(a, 3, "rgl_node_raytrace_configure_non_hits");
		     rgl_node_raytrace_configure_non_hits(lookup(st.nodes, a[0], "node"), asFloat(a[1]), asFloat(a[2]));
	     }},
	    {"rgl_node_raytrace_configure_mask", [](const Args& a, State& st) {
		     need(a, 3, "rgl_node_raytrace_configure_mask");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_raytrace_configure_mask(lookup(st.nodes, a[0], "node"), st.ptr<int8_t>(a[1], c), c);
	     }},
	    {"rgl_node_raytrace_configure_beam_divergence", [](const Args& a, State& st) {
		     need(a, 3, "rgl_node_raytrace_configure_beam_divergence");
		     rgl_node_raytrace_configure_beam_divergence(lookup(st.nodes, a[0], "node"), asFloat(a[1]), asFloat(a[2]));
	     }},
	    {"rgl_node_raytrace_configure_default_intensity", [](const Args& a, State& st) {
		     need(a, 2, "rgl_node_raytrace_configure_default_intensity");
		     rgl_node_raytrace_configure_default_intensity(lookup(st.nodes, a[0], "node"), asFloat(a[1]));
	     }},
	    {"rgl_node_raytrace_configure_return_mode", [](const Args& a, State& st) {
		     need(a, 2, "rgl_node_raytrace_configure_return_mode");
		     rgl_node_raytrace_configure_return_mode(lookup(st.nodes, a[0], "node"),
		                                             static_cast<rgl_return_mode_t>(asInt(a[1])));
	     }},
	    {"rgl_node_points_format", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_points_format");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_points_format(n, st.ptr<rgl_field_t>(a[1], c), c);
	     })},
	    {"rgl_node_points_yield", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_points_yield");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_points_yield(n, st.ptr<rgl_field_t>(a[1], c), c);
	     })},
	    {"rgl_node_points_compact_by_field", nodeCall([](rgl_node_t* n, const Args& a, State&) {
		     need(a, 2, "rgl_node_points_compact_by_field");
		     rgl_node_points_compact_by_field(n, static_cast<rgl_field_t>(asInt(a[1])));
	     })},
	    {"rgl_node_points_spatial_merge", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_points_spatial_merge");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_points_spatial_merge(n, st.ptr<rgl_field_t>(a[1], c), c);
	     })},
	    {"rgl_node_points_temporal_merge", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_points_temporal_merge");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_points_temporal_merge(n, st.ptr<rgl_field_t>(a[1], c), c);
	     })},
	    {"rgl_node_points_from_array", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 5, "rgl_node_points_from_array");
		     const int32_t fc = static_cast<int32_t>(asInt(a[4]));
		     rgl_node_points_from_array(n, st.ptr<void>(a[1]), static_cast<int32_t>(asInt(a[2])),
		                                st.ptr<rgl_field_t>(a[3], fc), fc);
	     })},
	    {"rgl_node_points_radar_postprocess", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 10, "rgl_node_points_radar_postprocess");
		     const int32_t c = static_cast<int32_t>(asInt(a[2]));
		     rgl_node_points_radar_postprocess(n, st.ptr<rgl_radar_scope_t>(a[1], c), c, asFloat(a[3]), asFloat(a[4]),
		                                       asFloat(a[5]), asFloat(a[6]), asFloat(a[7]), asFloat(a[8]),
		                                       asFloat(a[9]));
	     })},
	    {"rgl_node_points_radar_track_objects", nodeCall([](rgl_node_t* n, const Args& a, State&) {
		     need(a, 8, "rgl_node_points_radar_track_objects");
		     rgl_node_points_radar_track_objects(n, asFloat(a[1]), asFloat(a[2]), asFloat(a[3]), asFloat(a[4]),
		                                         asFloat(a[5]), asFloat(a[6]), asFloat(a[7]));
	     })},
	    {"rgl_node_points_filter_ground", nodeCall([](rgl_node_t* n, const Args& a, State& st) {
		     need(a, 3, "rgl_node_points_filter_ground");
		     rgl_node_points_filter_ground(n, st.ptr<rgl_vec3f>(a[1]), asFloat(a[2]));
	     })},
	    {"rgl_node_gaussian_noise_angular_ray", nodeCall([](rgl_node_t* n, const Args& a, State&) {
		     need(a, 4, "rgl_node_gaussian_noise_angular_ray");
		     rgl_node_gaussian_noise_angular_ray(n, asFloat(a[1]), asFloat(a[2]), static_cast<rgl_axis_t>(asInt(a[3])));
	     })},
	    {"rgl_node_gaussian_noise_angular_hitpoint", nodeCall([](rgl_node_t* n, const Args& a, State&) {
		     need(a, 4, "rgl_node_gaussian_noise_angular_hitpoint");
		     rgl_node_gaussian_noise_angular_hitpoint(n, asFloat(a[1]), asFloat(a[2]),
		                                              static_cast<rgl_axis_t>(asInt(a[3])));
	     })},
	    {"rgl_node_gaussian_noise_distance", nodeCall([](rgl_node_t* n, const Args& a, State&) {
		     need(a, 4, "rgl_node_gaussian_noise_distance");
		     rgl_node_gaussian_noise_distance(n, asFloat(a[1]), asFloat(a[2]), asFloat(a[3]));
	     })},
	};
	return table;
}

} // namespace

// ------------------------------------------------------------------------------------------------ player
void play(const std::string& path)
{
	const std::string yamlPath = path + ".yaml", binPath = path + ".bin";
	State st(binPath);
	std::ifstream in(yamlPath);
	if (!in) {
		throw InvalidFilePath("TAPE: could not open yaml file: '" + yamlPath + "'");
	}
	std::stringstream text;
	text << in.rdbuf();
	const std::vector<Call> calls = Parser(text.str()).parse();

	// src/tape/TapePlayer.cpp:38-53: the first version record must not be older than the format.
	const Call* version = nullptr;
	for (const Call& c : calls) {
		if (c.fn == "rgl_get_version_info") {
			version = &c;
			break;
		}
	}
	if (version == nullptr || version->args.size() < 3) {
		throw TapeError("Unsupported Tape format: Missing version record in the Tape");
	}
	const int64_t linear = 1000000 * asInt(version->args[0]) + 1000 * asInt(version->args[1]) + asInt(version->args[2]);
	if (linear < 1000000 * kFormatMajor + 1000 * kFormatMinor + kFormatPatch) {
		throw TapeError("Unsupported Tape Format: Tape version is too old. Required version: " +
		                std::to_string(kFormatMajor) + "." + std::to_string(kFormatMinor) + "." +
		                std::to_string(kFormatPatch) + ".");
	}
	for (const Call& c : calls) {
		if (!players().count(c.fn)) {
			throw TapeError("unknown function to play: " + c.fn);
		}
	}
	// src/tape/TapePlayer.cpp:114-126: each call at (approximately) its recorded time after the start.
	const int64_t start = nowNs();
	for (const Call& c : calls) {
		const int64_t wait = c.t - (nowNs() - start);
		if (wait > 0) {
			std::this_thread::sleep_for(std::chrono::nanoseconds(wait));
		}
		players().at(c.fn)(c.args, st);
	}
}

} // namespace rgl::tape
