// Minimal JSON reader for the drop-in's input.json (objects, arrays,
// strings, numbers, true/false/null).  at() throws std::out_of_range on a
// missing key, which is how the reference reacts to a missing required key
// (src/utils.h:82, src/kmeans.cpp:21-22).
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace thc {

class Json {
 public:
  enum Kind { Null, Bool, Number, String, Array, Object };
  Kind kind = Null;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  static Json parse(const std::string& text) {
    size_t p = 0;
    Json v = parse_value(text, p);
    skip(text, p);
    if (p != text.size()) throw std::runtime_error("json: trailing characters");
    return v;
  }

  bool contains(const std::string& key) const { return kind == Object && obj.count(key) > 0; }
  const Json& at(const std::string& key) const {
    if (kind != Object) throw std::out_of_range("json: not an object (key " + key + ")");
    auto it = obj.find(key);
    if (it == obj.end()) throw std::out_of_range("json: key not found: " + key);
    return it->second;
  }
  int as_int() const {
    if (kind != Number) throw std::runtime_error("json: not a number");
    return (int)num;
  }
  unsigned int as_uint() const { return (unsigned int)(long long)num; }
  double as_double() const {
    if (kind != Number) throw std::runtime_error("json: not a number");
    return num;
  }
  bool as_bool() const {
    if (kind != Bool) throw std::runtime_error("json: not a boolean");
    return b;
  }
  const std::string& as_string() const {
    if (kind != String) throw std::runtime_error("json: not a string");
    return str;
  }

 private:
  static void skip(const std::string& t, size_t& p) {
    while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
  }
  static Json parse_value(const std::string& t, size_t& p) {
    skip(t, p);
    if (p >= t.size()) throw std::runtime_error("json: unexpected end");
    const char c = t[p];
    Json v;
    if (c == '{') {
      v.kind = Object;
      ++p;
      skip(t, p);
      if (t[p] == '}') { ++p; return v; }
      for (;;) {
        skip(t, p);
        std::string key = parse_string(t, p);
        skip(t, p);
        if (t[p] != ':') throw std::runtime_error("json: expected ':'");
        ++p;
        v.obj[key] = parse_value(t, p);
        skip(t, p);
        if (t[p] == ',') { ++p; continue; }
        if (t[p] == '}') { ++p; return v; }
        throw std::runtime_error("json: expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Array;
      ++p;
      skip(t, p);
      if (t[p] == ']') { ++p; return v; }
      for (;;) {
        v.arr.push_back(parse_value(t, p));
        skip(t, p);
        if (t[p] == ',') { ++p; continue; }
        if (t[p] == ']') { ++p; return v; }
        throw std::runtime_error("json: expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = String;
      v.str = parse_string(t, p);
      return v;
    }
    if (t.compare(p, 4, "true") == 0) { v.kind = Bool; v.b = true; p += 4; return v; }
    if (t.compare(p, 5, "false") == 0) { v.kind = Bool; v.b = false; p += 5; return v; }
    if (t.compare(p, 4, "null") == 0) { p += 4; return v; }
    char* end = nullptr;
    v.num = std::strtod(t.c_str() + p, &end);
    if (end == t.c_str() + p) throw std::runtime_error("json: bad value");
    v.kind = Number;
    p = end - t.c_str();
    return v;
  }
  static std::string parse_string(const std::string& t, size_t& p) {
    if (t[p] != '"') throw std::runtime_error("json: expected string");
    ++p;
    std::string s;
    while (p < t.size() && t[p] != '"') {
      if (t[p] == '\\' && p + 1 < t.size()) {
        const char e = t[++p];
        s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
      } else {
        s += t[p];
      }
      ++p;
    }
    ++p;
    return s;
  }
};

}  // namespace thc
