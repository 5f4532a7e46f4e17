// Minimal JSON document model and parser for renderer configurations
// (RendererConfig / CameraConfig / *RendererParams of the reference,
// include/hydrant/core/renderer.schema.hpp, core/scene.schema.hpp,
// ui/*.schema.hpp).  Numbers are doubles; objects keep insertion order.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hyd {
namespace json {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Value {
 public:
  enum class Kind { Null, Bool, Number, String, Array, Object };

  Value() = default;
  static Value boolean(bool b) { Value v; v.kind_ = Kind::Bool; v.b_ = b; return v; }
  static Value number(double d) { Value v; v.kind_ = Kind::Number; v.d_ = d; return v; }
  static Value string(std::string s) { Value v; v.kind_ = Kind::String; v.s_ = std::move(s); return v; }
  static Value array() { Value v; v.kind_ = Kind::Array; return v; }
  static Value object() { Value v; v.kind_ = Kind::Object; return v; }

  Kind kind() const { return kind_; }
  bool is_null() const { return kind_ == Kind::Null; }
  bool is_number() const { return kind_ == Kind::Number; }
  bool is_string() const { return kind_ == Kind::String; }
  bool is_array() const { return kind_ == Kind::Array; }
  bool is_object() const { return kind_ == Kind::Object; }

  double as_number(const char *what) const {
    if (kind_ == Kind::Number) return d_;
    if (kind_ == Kind::Bool) return b_ ? 1.0 : 0.0;
    throw Error(std::string(what) + ": expected a number");
  }
  bool as_bool(const char *what) const {
    if (kind_ == Kind::Bool) return b_;
    if (kind_ == Kind::Number) return d_ != 0.0;
    throw Error(std::string(what) + ": expected a boolean");
  }
  const std::string &as_string(const char *what) const {
    if (kind_ != Kind::String) throw Error(std::string(what) + ": expected a string");
    return s_;
  }
  const std::vector<Value> &items(const char *what) const {
    if (kind_ != Kind::Array) throw Error(std::string(what) + ": expected an array");
    return a_;
  }
  std::vector<Value> &mutable_items() { return a_; }
  // object member or nullptr
  const Value *find(const std::string &key) const {
    if (kind_ != Kind::Object) return nullptr;
    for (const auto &kv : o_)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  const std::vector<std::pair<std::string, Value>> &members() const { return o_; }
  void set(std::string key, Value v) {
    for (auto &kv : o_)
      if (kv.first == key) { kv.second = std::move(v); return; }
    o_.emplace_back(std::move(key), std::move(v));
  }
  void push(Value v) { a_.push_back(std::move(v)); }

 private:
  Kind kind_ = Kind::Null;
  bool b_ = false;
  double d_ = 0.0;
  std::string s_;
  std::vector<Value> a_;
  std::vector<std::pair<std::string, Value>> o_;
};

class Parser {
 public:
  explicit Parser(const std::string &text) : t_(text) {}

  Value parse() {
    Value v = value();
    ws();
    if (i_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &t_;
  size_t i_ = 0;
  int depth_ = 0;

  [[noreturn]] void fail(const char *msg) const {
    throw Error("JSON parse error at offset " + std::to_string(i_) + ": " + msg);
  }
  void ws() {
    while (i_ < t_.size() && (t_[i_] == ' ' || t_[i_] == '\t' || t_[i_] == '\n' || t_[i_] == '\r')) ++i_;
  }
  bool lit(const char *w) {
    size_t n = 0;
    while (w[n]) ++n;
    if (t_.compare(i_, n, w) == 0) { i_ += n; return true; }
    return false;
  }
  Value value() {
    ws();
    if (i_ >= t_.size()) fail("unexpected end of input");
    if (++depth_ > 256) fail("nesting too deep");
    Value v;
    const char c = t_[i_];
    if (c == '{') v = object();
    else if (c == '[') v = array();
    else if (c == '"') v = Value::string(str());
    else if (lit("true")) v = Value::boolean(true);
    else if (lit("false")) v = Value::boolean(false);
    else if (lit("null")) v = Value();
    else v = Value::number(num());
    --depth_;
    return v;
  }
  Value object() {
    Value o = Value::object();
    ++i_;
    ws();
    if (i_ < t_.size() && t_[i_] == '}') { ++i_; return o; }
    for (;;) {
      ws();
      if (i_ >= t_.size() || t_[i_] != '"') fail("expected a member name");
      std::string k = str();
      ws();
      if (i_ >= t_.size() || t_[i_] != ':') fail("expected ':'");
      ++i_;
      o.set(std::move(k), value());
      ws();
      if (i_ < t_.size() && t_[i_] == ',') { ++i_; continue; }
      if (i_ < t_.size() && t_[i_] == '}') { ++i_; return o; }
      fail("expected ',' or '}'");
    }
  }
  Value array() {
    Value a = Value::array();
    ++i_;
    ws();
    if (i_ < t_.size() && t_[i_] == ']') { ++i_; return a; }
    for (;;) {
      a.push(value());
      ws();
      if (i_ < t_.size() && t_[i_] == ',') { ++i_; continue; }
      if (i_ < t_.size() && t_[i_] == ']') { ++i_; return a; }
      fail("expected ',' or ']'");
    }
  }
  static void utf8(std::string &out, uint32_t cp) {
    if (cp < 0x80) out += char(cp);
    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
      out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F));
    } else {
      out += char(0xF0 | (cp >> 18)); out += char(0x80 | ((cp >> 12) & 0x3F));
      out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (i_ + 4 > t_.size()) fail("short \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = t_[i_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
      else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string str() {
    ++i_;   // opening quote
    std::string out;
    for (;;) {
      if (i_ >= t_.size()) fail("unterminated string");
      const char c = t_[i_++];
      if (c == '"') return out;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') { out += c; continue; }
      if (i_ >= t_.size()) fail("unterminated escape");
      const char e = t_[i_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && i_ + 1 < t_.size() && t_[i_] == '\\' && t_[i_ + 1] == 'u') {
            i_ += 2;
            const uint32_t lo = hex4();
            if (lo < 0xDC00 || lo > 0xDFFF) fail("bad surrogate pair");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
  double num() {
    const size_t s = i_;
    if (i_ < t_.size() && (t_[i_] == '-' || t_[i_] == '+')) ++i_;
    bool digits = false;
    while (i_ < t_.size() && ((t_[i_] >= '0' && t_[i_] <= '9') || t_[i_] == '.' || t_[i_] == 'e' ||
                              t_[i_] == 'E' || ((t_[i_] == '-' || t_[i_] == '+') &&
                                                (t_[i_ - 1] == 'e' || t_[i_ - 1] == 'E')))) {
      digits |= (t_[i_] >= '0' && t_[i_] <= '9');
      ++i_;
    }
    if (!digits) fail("expected a value");
    const std::string tok = t_.substr(s, i_ - s);
    char *end = nullptr;
    const double d = std::strtod(tok.c_str(), &end);
    if (!end || *end) fail("malformed number");
    return d;
  }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace hyd
