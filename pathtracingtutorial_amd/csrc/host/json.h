// Minimal JSON reader for scene files.
//
// Objects keep their keys sorted (std::map), which is the iteration order the
// reference's scene loader sees for "materials" (nlohmann::json's default
// object type is a std::map, Part_2/src/scene/scene_loader.cpp:112-122), so
// material indices come out the same.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace ptamd {

struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    bool is_number() const { return type == Number; }
    bool is_string() const { return type == String; }
    bool is_array() const { return type == Array; }
    bool is_object() const { return type == Object; }
    bool contains(const std::string& k) const { return type == Object && obj.count(k) != 0; }
    const Json& operator[](const std::string& k) const;
    const Json& operator[](size_t i) const;
    size_t size() const { return type == Array ? arr.size() : (type == Object ? obj.size() : 0); }
    float as_float() const;
    int as_int() const;
    bool as_bool() const;
};

// Parses text; on failure returns false and sets *err with the byte offset.
bool json_parse(const std::string& text, Json& out, std::string* err);

}  // namespace ptamd
