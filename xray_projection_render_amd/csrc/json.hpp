// json.hpp - small JSON document model and parser for scene and
// deformation descriptions.
//
// Numbers remember whether their literal was an integer so the loader can
// apply the reference's decoder semantics: Go's encoding/json turns every
// number into float64, yaml.v3 keeps integer literals as int.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace xrp {
namespace json {

struct Value;
using Array = std::vector<Value>;
using Object = std::vector<std::pair<std::string, Value>>;

struct Value {
  enum Type { Null, Bool, Int, Float, String, Arr, Obj } type = Null;
  bool b = false;
  int64_t i = 0;
  double f = 0.0;
  std::string s;
  std::shared_ptr<Array> a;
  std::shared_ptr<Object> o;

  bool is_null() const { return type == Null; }
  bool is_number() const { return type == Int || type == Float; }
  double number() const { return type == Int ? (double)i : f; }
  const Value* get(const std::string& key) const {
    if (type != Obj) return nullptr;
    for (auto& kv : *o)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  bool has(const std::string& key) const { return get(key) != nullptr; }
};

class ParseError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Parser {
 public:
  explicit Parser(const std::string& text) : t_(text) {}

  Value parse() {
    Value v = value();
    ws();
    if (p_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& t_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) {
    throw ParseError(std::string("JSON parse error at offset ") +
                     std::to_string(p_) + ": " + what);
  }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' ||
                              t_[p_] == '\r' || t_[p_] == '\t'))
      ++p_;
  }
  bool lit(const char* w) {
    size_t n = 0;
    while (w[n]) ++n;
    if (t_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Value value() {
    ws();
    if (p_ >= t_.size()) fail("unexpected end");
    char c = t_[p_];
    Value v;
    if (c == '{') {
      ++p_;
      v.type = Value::Obj;
      v.o = std::make_shared<Object>();
      ws();
      if (p_ < t_.size() && t_[p_] == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= t_.size() || t_[p_] != '"') fail("expected key");
        std::string k = str();
        ws();
        if (p_ >= t_.size() || t_[p_] != ':') fail("expected ':'");
        ++p_;
        Value item = value();
        bool replaced = false;  // duplicate keys: last one wins (as Go)
        for (auto& kv : *v.o)
          if (kv.first == k) {
            kv.second = item;
            replaced = true;
          }
        if (!replaced) v.o->emplace_back(k, item);
        ws();
        if (p_ < t_.size() && t_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < t_.size() && t_[p_] == '}') {
          ++p_;
          return v;
        }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p_;
      v.type = Value::Arr;
      v.a = std::make_shared<Array>();
      ws();
      if (p_ < t_.size() && t_[p_] == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.a->push_back(value());
        ws();
        if (p_ < t_.size() && t_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < t_.size() && t_[p_] == ']') {
          ++p_;
          return v;
        }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.type = Value::String;
      v.s = str();
      return v;
    }
    if (lit("true")) {
      v.type = Value::Bool;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.type = Value::Bool;
      return v;
    }
    if (lit("null")) return v;
    return number();
  }
  std::string str() {
    ++p_;  // opening quote
    std::string out;
    while (p_ < t_.size() && t_[p_] != '"') {
      char c = t_[p_++];
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= t_.size()) fail("bad escape");
      char e = t_[p_++];
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
          if (p_ + 4 > t_.size()) fail("bad \\u escape");
          unsigned cp = std::stoul(t_.substr(p_, 4), nullptr, 16);
          p_ += 4;
          if (cp < 0x80) {
            out += (char)cp;
          } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
          } else {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
          }
          break;
        }
        default: fail("bad escape");
      }
    }
    if (p_ >= t_.size()) fail("unterminated string");
    ++p_;
    return out;
  }
  Value number() {
    size_t s = p_;
    bool is_float = false;
    if (p_ < t_.size() && (t_[p_] == '-' || t_[p_] == '+')) ++p_;
    while (p_ < t_.size()) {
      char c = t_[p_];
      if (c >= '0' && c <= '9') {
        ++p_;
      } else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') {
        is_float = true;
        ++p_;
      } else {
        break;
      }
    }
    if (p_ == s) fail("unexpected character");
    std::string tok = t_.substr(s, p_ - s);
    Value v;
    try {
      if (!is_float) {
        v.type = Value::Int;
        v.i = std::stoll(tok);
        v.f = (double)v.i;
      } else {
        v.type = Value::Float;
        v.f = std::stod(tok);
      }
    } catch (const std::exception&) {
      fail("bad number");
    }
    return v;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace xrp
