// yaml_lite.hpp - the YAML subset scene and deformation files use, decoded
// into the json::Value model with yaml.v3's scalar typing (integer literals
// stay Int, decimals/exponents are Float, true/false Bool, null/~ Null).
//
// Supported: block mappings and sequences by indentation (a sequence may sit
// at its parent key's indentation), "- key: value" items, flow sequences
// [a, b] and flow mappings {k: v}, quoted and plain scalars, # comments.
#pragma once
#include <cctype>
#include <string>
#include <vector>

#include "json.hpp"

namespace xrp {
namespace yaml {

using json::Value;

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

struct Line {
  int indent;
  std::string text;
  int number;
};

inline std::string strip_comment(const std::string& s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    if (c == '"' && !sq) dq = !dq;
    if (c == '#' && !sq && !dq && (i == 0 || std::isspace((unsigned char)s[i - 1])))
      return s.substr(0, i);
  }
  return s;
}

inline std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

inline Value scalar(const std::string& raw);

inline std::vector<std::string> split_flow(const std::string& s) {
  std::vector<std::string> out;
  int depth = 0;
  bool sq = false, dq = false;
  std::string cur;
  for (char c : s) {
    if (c == '\'' && !dq) sq = !sq;
    if (c == '"' && !sq) dq = !dq;
    if (!sq && !dq) {
      if (c == '[' || c == '{') ++depth;
      if (c == ']' || c == '}') --depth;
      if (c == ',' && depth == 0) {
        out.push_back(trim(cur));
        cur.clear();
        continue;
      }
    }
    cur += c;
  }
  if (!trim(cur).empty()) out.push_back(trim(cur));
  return out;
}

// Position of the ':' ending a mapping key, or npos.
inline size_t key_colon(const std::string& s) {
  bool sq = false, dq = false;
  int depth = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    if (c == '"' && !sq) dq = !dq;
    if (sq || dq) continue;
    if (c == '[' || c == '{') ++depth;
    if (c == ']' || c == '}') --depth;
    if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
  }
  return std::string::npos;
}

inline std::string unquote(const std::string& k) {
  if (k.size() >= 2 && ((k.front() == '"' && k.back() == '"') ||
                        (k.front() == '\'' && k.back() == '\'')))
    return k.substr(1, k.size() - 2);
  return k;
}

inline Value scalar(const std::string& raw) {
  std::string s = trim(raw);
  Value v;
  if (s.empty() || s == "~" || s == "null" || s == "Null" || s == "NULL") return v;
  if (s.front() == '[') {
    if (s.back() != ']') throw Error("unterminated flow sequence: " + s);
    v.type = Value::Arr;
    v.a = std::make_shared<json::Array>();
    for (auto& item : split_flow(s.substr(1, s.size() - 2))) v.a->push_back(scalar(item));
    return v;
  }
  if (s.front() == '{') {
    if (s.back() != '}') throw Error("unterminated flow mapping: " + s);
    v.type = Value::Obj;
    v.o = std::make_shared<json::Object>();
    for (auto& item : split_flow(s.substr(1, s.size() - 2))) {
      size_t c = key_colon(item);
      if (c == std::string::npos) throw Error("bad flow mapping entry: " + item);
      v.o->emplace_back(unquote(trim(item.substr(0, c))), scalar(item.substr(c + 1)));
    }
    return v;
  }
  if (s.front() == '"' || s.front() == '\'') {
    v.type = Value::String;
    v.s = unquote(s);
    return v;
  }
  if (s == "true" || s == "True" || s == "TRUE" || s == "false" || s == "False" ||
      s == "FALSE") {
    v.type = Value::Bool;
    v.b = s[0] == 't' || s[0] == 'T';
    return v;
  }
  // yaml.v3 core schema numbers
  size_t i = 0;
  if (i < s.size() && (s[i] == '-' || s[i] == '+')) ++i;
  bool digits = false, dot = false, exp = false, ok = i < s.size();
  for (size_t k = i; k < s.size() && ok; ++k) {
    char c = s[k];
    if (std::isdigit((unsigned char)c)) {
      digits = true;
    } else if (c == '.' && !dot && !exp) {
      dot = true;
    } else if ((c == 'e' || c == 'E') && digits && !exp) {
      exp = true;
      if (k + 1 < s.size() && (s[k + 1] == '-' || s[k + 1] == '+')) ++k;
      if (k + 1 >= s.size()) ok = false;
    } else {
      ok = false;
    }
  }
  if (ok && digits) {
    try {
      if (!dot && !exp) {
        v.type = Value::Int;
        v.i = std::stoll(s);
        v.f = (double)v.i;
      } else {
        v.type = Value::Float;
        v.f = std::stod(s);
      }
      return v;
    } catch (const std::exception&) {
    }
  }
  std::string low = s;
  for (auto& c : low) c = (char)std::tolower((unsigned char)c);
  if (low == ".inf" || low == "+.inf" || low == "-.inf" || low == ".nan") {
    v.type = Value::Float;
    v.f = low == ".nan" ? NAN : (low[0] == '-' ? -INFINITY : INFINITY);
    return v;
  }
  v.type = Value::String;
  v.s = s;
  return v;
}

class Parser {
 public:
  explicit Parser(const std::string& text) {
    int n = 0;
    size_t p = 0;
    while (p <= text.size()) {
      size_t e = text.find('\n', p);
      if (e == std::string::npos) e = text.size();
      std::string line = text.substr(p, e - p);
      ++n;
      if (!line.empty() && line.back() == '\r') line.pop_back();
      std::string body = strip_comment(line);
      if (trim(body) == "---") body.clear();
      if (!trim(body).empty()) {
        int ind = 0;
        while (ind < (int)body.size() && body[ind] == ' ') ++ind;
        if (ind < (int)body.size() && body[ind] == '\t')
          throw Error("tab indentation at line " + std::to_string(n));
        lines_.push_back({ind, trim(body), n});
      }
      p = e + 1;
    }
  }

  Value parse() {
    if (lines_.empty()) return Value();
    size_t i = 0;
    Value v = block(i, lines_[0].indent);
    if (i != lines_.size())
      throw Error("unexpected indentation at line " + std::to_string(lines_[i].number));
    return v;
  }

 private:
  std::vector<Line> lines_;

  static bool is_item(const std::string& t) {
    return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' ');
  }

  Value block(size_t& i, int indent) {
    if (is_item(lines_[i].text)) return sequence(i, indent);
    if (key_colon(lines_[i].text) != std::string::npos) return mapping(i, indent);
    Value v = scalar(lines_[i].text);
    ++i;
    return v;
  }

  Value value_after(size_t& i, int parent_indent, const std::string& rest,
                    bool seq_same_indent) {
    if (!trim(rest).empty()) return scalar(rest);
    if (i < lines_.size() && lines_[i].indent > parent_indent)
      return block(i, lines_[i].indent);
    if (seq_same_indent && i < lines_.size() && lines_[i].indent == parent_indent &&
        is_item(lines_[i].text))
      return sequence(i, parent_indent);
    return Value();
  }

  Value mapping(size_t& i, int indent) {
    Value v;
    v.type = Value::Obj;
    v.o = std::make_shared<json::Object>();
    while (i < lines_.size() && lines_[i].indent == indent && !is_item(lines_[i].text)) {
      const std::string t = lines_[i].text;
      size_t c = key_colon(t);
      if (c == std::string::npos)
        throw Error("expected 'key: value' at line " + std::to_string(lines_[i].number));
      std::string key = unquote(trim(t.substr(0, c)));
      std::string rest = t.substr(c + 1);
      ++i;
      Value item = value_after(i, indent, rest, true);
      bool replaced = false;
      for (auto& kv : *v.o)
        if (kv.first == key) {
          kv.second = item;
          replaced = true;
        }
      if (!replaced) v.o->emplace_back(key, item);
    }
    return v;
  }

  Value sequence(size_t& i, int indent) {
    Value v;
    v.type = Value::Arr;
    v.a = std::make_shared<json::Array>();
    while (i < lines_.size() && lines_[i].indent == indent && is_item(lines_[i].text)) {
      std::string rest = lines_[i].text.size() > 1 ? lines_[i].text.substr(2) : "";
      std::string r = trim(rest);
      if (r.empty()) {
        ++i;
        v.a->push_back(value_after(i, indent, "", false));
      } else if (key_colon(r) != std::string::npos && r.front() != '[' && r.front() != '{' &&
                 r.front() != '"' && r.front() != '\'') {
        // "- key: value": a mapping whose keys sit at indent + 2
        int inner = indent + 2 + (int)rest.find_first_not_of(' ');
        lines_[i].indent = inner;
        lines_[i].text = r;
        v.a->push_back(mapping(i, inner));
      } else {
        ++i;
        v.a->push_back(scalar(r));
      }
    }
    return v;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace yaml
}  // namespace xrp
