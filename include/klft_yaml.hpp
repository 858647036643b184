// klft_yaml.hpp -- the YAML subset read by klft's input files.
//
// yaml-cpp (the reference's parser, InputParser.hpp:25) is not part of this
// build; this header parses what klft inputs use: nested block mappings,
// block sequences ("- item"), flow sequences ("[2, 2]", nested), plain and
// quoted scalars, booleans and comments.  Conversions follow yaml-cpp's
// Node::as<T>() / as<T>(fallback): as<T>(fallback) returns the fallback when
// the node is missing, null or does not convert; as<T>() throws YamlError.
#pragma once
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace klft_yaml {

struct YamlError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Node {
  enum Kind { Undefined, Null, Scalar, Seq, Map };
  Kind kind = Undefined;
  std::string text;
  std::vector<Node> items;
  std::vector<std::pair<std::string, Node>> fields;

  static const Node &undefined() {
    static const Node u;
    return u;
  }
  // yaml-cpp: a node is truthy when it is defined
  explicit operator bool() const { return kind != Undefined; }
  bool IsSequence() const { return kind == Seq; }
  bool IsMap() const { return kind == Map; }
  size_t size() const { return kind == Seq ? items.size() : (kind == Map ? fields.size() : 0); }

  const Node &operator[](const std::string &key) const {
    if (kind != Map) return undefined();
    for (const auto &f : fields)
      if (f.first == key) return f.second;
    return undefined();
  }
  const Node &operator[](const char *key) const { return (*this)[std::string(key)]; }
  const Node &operator[](size_t i) const {
    if (kind != Seq || i >= items.size()) return undefined();
    return items[i];
  }
  const Node &operator[](int i) const { return (*this)[(size_t)i]; }
  std::vector<Node>::const_iterator begin() const { return items.begin(); }
  std::vector<Node>::const_iterator end() const { return items.end(); }

  template <typename T>
  bool decode(T &out) const;

  template <typename T>
  T as() const {
    T v{};
    if (!decode(v)) throw YamlError("bad conversion");
    return v;
  }
  template <typename T, typename S>
  T as(const S &fallback) const {
    T v{};
    if (kind != Scalar || !decode(v)) return static_cast<T>(fallback);
    return v;
  }
};

namespace detail {

inline std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

inline bool parse_int(const std::string &t, long long &v) {
  if (t.empty()) return false;
  size_t i = 0;
  bool neg = false;
  if (t[i] == '+' || t[i] == '-') neg = t[i++] == '-';
  if (i >= t.size()) return false;
  int base = 10;
  if (t.size() - i > 2 && t[i] == '0' && (t[i + 1] == 'x' || t[i + 1] == 'X')) {
    base = 16;
    i += 2;
  } else if (t.size() - i > 2 && t[i] == '0' && t[i + 1] == 'o') {
    base = 8;
    i += 2;
  }
  errno = 0;
  char *end = nullptr;
  const unsigned long long u = std::strtoull(t.c_str() + i, &end, base);
  if (errno || end == t.c_str() + i || *end != '\0') return false;
  if (!std::isalnum((unsigned char)t[i])) return false;
  v = neg ? -(long long)u : (long long)u;
  return true;
}

inline bool parse_double(const std::string &t, double &v) {
  if (t == ".inf" || t == ".Inf" || t == ".INF" || t == "+.inf" || t == "+.Inf" || t == "+.INF") {
    v = HUGE_VAL;
    return true;
  }
  if (t == "-.inf" || t == "-.Inf" || t == "-.INF") {
    v = -HUGE_VAL;
    return true;
  }
  if (t == ".nan" || t == ".NaN" || t == ".NAN") {
    v = std::nan("");
    return true;
  }
  if (t.empty()) return false;
  char *end = nullptr;
  errno = 0;
  v = std::strtod(t.c_str(), &end);
  return end != t.c_str() && *end == '\0' && !std::isspace((unsigned char)t[0]);
}

inline bool parse_bool(const std::string &t, bool &v) {
  static const char *yes[] = {"y", "Y", "yes", "Yes", "YES", "true", "True", "TRUE", "on", "On", "ON"};
  static const char *no[] = {"n", "N", "no", "No", "NO", "false", "False", "FALSE", "off", "Off", "OFF"};
  for (const char *s : yes)
    if (t == s) {
      v = true;
      return true;
    }
  for (const char *s : no)
    if (t == s) {
      v = false;
      return true;
    }
  return false;
}

struct Line {
  int indent;
  std::string text;
  int lineno;
};

// remove a comment ('#' at the start or after whitespace, outside quotes)
inline std::string strip_comment(const std::string &s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') {
      q = c;
      continue;
    }
    if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

class Parser {
 public:
  explicit Parser(const std::string &src) {
    std::istringstream in(src);
    std::string raw;
    int n = 0;
    while (std::getline(in, raw)) {
      ++n;
      if (!raw.empty() && raw.back() == '\r') raw.pop_back();
      std::string s = strip_comment(raw);
      size_t ind = 0;
      while (ind < s.size() && s[ind] == ' ') ++ind;
      if (ind < s.size() && s[ind] == '\t') throw YamlError("tab indentation at line " + std::to_string(n));
      std::string t = trim(s);
      if (t.empty() || t == "---" || t == "...") continue;
      lines_.push_back(Line{(int)ind, t, n});
    }
  }

  Node parse() {
    if (lines_.empty()) {
      Node n;
      n.kind = Node::Null;
      return n;
    }
    size_t i = 0;
    Node n = block(i, lines_[0].indent);
    if (i != lines_.size()) throw YamlError("unexpected content at line " + std::to_string(lines_[i].lineno));
    return n;
  }

 private:
  std::vector<Line> lines_;

  static bool is_item(const std::string &t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

  // position of the ':' separating a mapping key, or npos
  static size_t key_colon(const std::string &t) {
    char q = 0;
    int depth = 0;
    for (size_t i = 0; i < t.size(); ++i) {
      const char c = t[i];
      if (q) {
        if (c == q) q = 0;
        continue;
      }
      if (c == '"' || c == '\'') q = c;
      else if (c == '[' || c == '{') ++depth;
      else if (c == ']' || c == '}') --depth;
      else if (c == ':' && depth == 0 && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
    }
    return std::string::npos;
  }

  static std::string unquote(const std::string &t) {
    if (t.size() >= 2 && ((t.front() == '"' && t.back() == '"') || (t.front() == '\'' && t.back() == '\'')))
      return t.substr(1, t.size() - 2);
    return t;
  }

  Node block(size_t &i, int indent) {
    if (is_item(lines_[i].text)) return seq(i, indent);
    return map(i, indent);
  }

  Node seq(size_t &i, int indent) {
    Node n;
    n.kind = Node::Seq;
    while (i < lines_.size() && lines_[i].indent == indent && is_item(lines_[i].text)) {
      const Line &l = lines_[i];
      std::string rest = l.text == "-" ? "" : trim(l.text.substr(2));
      ++i;
      if (rest.empty()) {
        if (i < lines_.size() && lines_[i].indent > indent)
          n.items.push_back(block(i, lines_[i].indent));
        else {
          Node z;
          z.kind = Node::Null;
          n.items.push_back(z);
        }
      } else if (key_colon(rest) != std::string::npos && rest[0] != '[' && rest[0] != '{') {
        // "- key: value" starts a mapping whose further keys are indented
        const int sub = indent + 2 + (int)(l.text.size() - 2 - trim(l.text.substr(2)).size());
        lines_.insert(lines_.begin() + i, Line{sub, rest, l.lineno});
        n.items.push_back(map(i, sub));
      } else {
        n.items.push_back(inline_value(rest, l.lineno));
      }
    }
    return n;
  }

  Node map(size_t &i, int indent) {
    Node n;
    n.kind = Node::Map;
    while (i < lines_.size() && lines_[i].indent == indent) {
      const Line &l = lines_[i];
      if (is_item(l.text)) throw YamlError("unexpected sequence item at line " + std::to_string(l.lineno));
      const size_t c = key_colon(l.text);
      if (c == std::string::npos) throw YamlError("expected 'key: value' at line " + std::to_string(l.lineno));
      const std::string key = unquote(trim(l.text.substr(0, c)));
      const std::string rest = trim(l.text.substr(c + 1));
      const int lineno = l.lineno;
      ++i;
      Node v;
      if (!rest.empty()) {
        v = inline_value(rest, lineno);
      } else if (i < lines_.size() && lines_[i].indent > indent) {
        v = block(i, lines_[i].indent);
      } else if (i < lines_.size() && lines_[i].indent == indent && is_item(lines_[i].text)) {
        v = seq(i, indent);
      } else {
        v.kind = Node::Null;
      }
      for (auto &f : n.fields)
        if (f.first == key) throw YamlError("duplicate key '" + key + "' at line " + std::to_string(lineno));
      n.fields.emplace_back(key, std::move(v));
    }
    if (i < lines_.size() && lines_[i].indent > indent)
      throw YamlError("bad indentation at line " + std::to_string(lines_[i].lineno));
    return n;
  }

  Node inline_value(const std::string &t, int lineno) {
    size_t p = 0;
    Node n = flow(t, p, lineno);
    while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
    if (p != t.size()) throw YamlError("trailing characters at line " + std::to_string(lineno));
    return n;
  }

  Node flow(const std::string &t, size_t &p, int lineno) {
    while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
    Node n;
    if (p < t.size() && t[p] == '[') {
      n.kind = Node::Seq;
      ++p;
      for (;;) {
        while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
        if (p >= t.size()) throw YamlError("unterminated '[' at line " + std::to_string(lineno));
        if (t[p] == ']') {
          ++p;
          break;
        }
        n.items.push_back(flow(t, p, lineno));
        while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
        if (p < t.size() && t[p] == ',') {
          ++p;
          continue;
        }
        if (p < t.size() && t[p] == ']') {
          ++p;
          break;
        }
        throw YamlError("expected ',' or ']' at line " + std::to_string(lineno));
      }
      return n;
    }
    if (p < t.size() && t[p] == '{') {
      n.kind = Node::Map;
      ++p;
      for (;;) {
        while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
        if (p >= t.size()) throw YamlError("unterminated '{' at line " + std::to_string(lineno));
        if (t[p] == '}') {
          ++p;
          break;
        }
        size_t c = t.find(':', p);
        if (c == std::string::npos) throw YamlError("expected ':' at line " + std::to_string(lineno));
        const std::string key = unquote(trim(t.substr(p, c - p)));
        p = c + 1;
        n.fields.emplace_back(key, flow(t, p, lineno));
        while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
        if (p < t.size() && t[p] == ',') {
          ++p;
          continue;
        }
        if (p < t.size() && t[p] == '}') {
          ++p;
          break;
        }
        throw YamlError("expected ',' or '}' at line " + std::to_string(lineno));
      }
      return n;
    }
    if (p < t.size() && (t[p] == '"' || t[p] == '\'')) {
      const char q = t[p++];
      std::string s;
      while (p < t.size() && t[p] != q) s += t[p++];
      if (p >= t.size()) throw YamlError("unterminated string at line " + std::to_string(lineno));
      ++p;
      n.kind = Node::Scalar;
      n.text = s;
      return n;
    }
    const size_t a = p;
    int depth = 0;
    while (p < t.size()) {
      const char c = t[p];
      if (depth == 0 && (c == ',' || c == ']' || c == '}')) break;
      ++p;
    }
    const std::string s = trim(t.substr(a, p - a));
    if (s == "~" || s == "null" || s == "Null" || s == "NULL" || s.empty()) {
      n.kind = Node::Null;
    } else {
      n.kind = Node::Scalar;
      n.text = s;
    }
    return n;
  }
};

}  // namespace detail

template <typename T>
bool Node::decode(T &out) const {
  if (kind != Scalar) return false;
  if constexpr (std::is_same<T, bool>::value) {
    return detail::parse_bool(text, out);
  } else if constexpr (std::is_same<T, std::string>::value) {
    out = text;
    return true;
  } else if constexpr (std::is_floating_point<T>::value) {
    double d;
    if (!detail::parse_double(text, d)) return false;
    out = (T)d;
    return true;
  } else if constexpr (std::is_integral<T>::value) {
    long long v;
    if (!detail::parse_int(text, v)) return false;
    if (std::is_unsigned<T>::value && v < 0) return false;
    out = (T)v;
    return (long long)out == v;
  } else {
    return false;
  }
}

inline Node Load(const std::string &src) { return detail::Parser(src).parse(); }

inline Node LoadFile(const std::string &path) {
  std::ifstream in(path);
  if (!in.is_open()) throw YamlError("bad file: " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  return Load(ss.str());
}

}  // namespace klft_yaml
