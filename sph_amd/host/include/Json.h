// Json.h -- minimal JSON value, reader and writer for the scene and frame
// formats (the reference reads scenes with nlohmann/json,
// src/SceneLoader.cpp:9-36; this is a small reader of its own).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sph {

class JsonError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Json {
public:
    enum Type { Null, Bool, Number, String, Array, Object };

    Json() : type_(Null) {}
    static Json number(double v) { Json j; j.type_ = Number; j.num_ = v; return j; }
    static Json boolean(bool v) { Json j; j.type_ = Bool; j.b_ = v; return j; }
    static Json string(const std::string& s) { Json j; j.type_ = String; j.str_ = s; return j; }
    static Json array() { Json j; j.type_ = Array; return j; }
    static Json object() { Json j; j.type_ = Object; return j; }

    static Json parse(const std::string& text);
    static Json parseFile(const std::string& path);

    Type type() const { return type_; }
    bool isNull() const { return type_ == Null; }
    bool isNumber() const { return type_ == Number; }
    bool isArray() const { return type_ == Array; }
    bool isObject() const { return type_ == Object; }

    // Typed access; throws JsonError naming `what` when the type is wrong
    // (the reference throws on a null -> number conversion of a missing key).
    double num(const char* what = "value") const;
    int integer(const char* what = "value") const;
    bool flag(const char* what = "value") const;
    const std::string& str(const char* what = "value") const;

    // Object access.  `at` throws on a missing key; `get` returns null.
    bool has(const std::string& key) const;
    const Json& at(const std::string& key) const;
    const Json& get(const std::string& key) const;
    Json& operator[](const std::string& key);   // inserts (object)

    // Array access.
    size_t size() const;
    const Json& at(size_t i) const;
    void push(const Json& v);

    std::string dump() const;

private:
    Type type_;
    double num_ = 0;
    bool b_ = false;
    std::string str_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;   // insertion order
    void dumpTo(std::string& out) const;
    friend class JsonReader;
};

}  // namespace sph
