// gwm_json.hpp -- JSON reader of the host mirror's prepared-case ingest.
//
// Mirrors the reader interface of include/gwm/ingest/json_value.hpp
// (gwm::ingest::json::Value, parse_text, parse_file, to_real_vector,
// to_string_map); the reference parses with the third-party nlohmann/json,
// which this repository does not use.  A recursive-descent parser over the
// whole text.  Arrays whose elements are all numbers -- the field payloads of
// the analysis and boundary files, millions of values -- are stored packed as
// doubles (8 bytes per value instead of a tree node per value) and expanded
// into Values only if as_array() is asked for.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <variant>
#include <vector>

#include "gwm/gwm_mirror.hpp"

namespace gwm::ingest::json {

class Value {
 public:
  using Object = std::unordered_map<std::string, Value>;
  using Array = std::vector<Value>;
  using Numbers = std::vector<double>;

  Value() = default;
  explicit Value(std::nullptr_t) : v_(nullptr) {}
  explicit Value(bool b) : v_(b) {}
  explicit Value(double d) : v_(d) {}
  explicit Value(std::string s) : v_(std::move(s)) {}
  explicit Value(Array a) : v_(std::move(a)) {}
  explicit Value(Object o) : v_(std::make_shared<const Object>(std::move(o))) {}
  static Value numbers(Numbers n) {
    Value v;
    v.v_ = std::move(n);
    return v;
  }

  bool is_null() const { return std::holds_alternative<std::nullptr_t>(v_); }
  bool is_bool() const { return std::holds_alternative<bool>(v_); }
  bool is_number() const { return std::holds_alternative<double>(v_); }
  bool is_string() const { return std::holds_alternative<std::string>(v_); }
  bool is_array() const {
    return std::holds_alternative<Array>(v_) || std::holds_alternative<Numbers>(v_);
  }
  bool is_object() const { return std::holds_alternative<ObjectPtr>(v_); }
  // Packed all-number array (extension).
  bool is_numbers() const { return std::holds_alternative<Numbers>(v_); }

  bool as_bool() const {
    require(is_bool(), "JSON value is not bool");
    return std::get<bool>(v_);
  }
  double as_number() const {
    require(is_number(), "JSON value is not number");
    return std::get<double>(v_);
  }
  int as_int() const {
    const double d = as_number();
    require(std::fabs(d - std::round(d)) < 1.0e-6, "JSON numeric value is not integral");
    return static_cast<int>(std::llround(d));
  }
  std::uint64_t as_uint64() const {
    const double d = as_number();
    require(d >= 0.0, "JSON numeric value is negative");
    require(std::fabs(d - std::round(d)) < 1.0e-6, "JSON numeric value is not integral");
    return static_cast<std::uint64_t>(std::llround(d));
  }
  const std::string& as_string() const {
    require(is_string(), "JSON value is not string");
    return std::get<std::string>(v_);
  }
  const Numbers& as_numbers() const {
    require(is_numbers(), "JSON value is not a number array");
    return std::get<Numbers>(v_);
  }
  const Array& as_array() const {
    if (const auto* n = std::get_if<Numbers>(&v_)) {
      if (!expanded_) {
        auto a = std::make_shared<Array>();
        a->reserve(n->size());
        for (double d : *n) a->emplace_back(d);
        expanded_ = std::move(a);
      }
      return *expanded_;
    }
    require(std::holds_alternative<Array>(v_), "JSON value is not array");
    return std::get<Array>(v_);
  }
  const Object& as_object() const {
    require(is_object(), "JSON value is not object");
    return *std::get<ObjectPtr>(v_);
  }
  const Value& at(const std::string& key) const {
    const auto& o = as_object();
    const auto it = o.find(key);
    require(it != o.end(), "Missing JSON object key: " + key);
    return it->second;
  }
  bool has(const std::string& key) const {
    return is_object() && as_object().count(key) != 0;
  }
  std::size_t size() const {
    if (const auto* n = std::get_if<Numbers>(&v_)) return n->size();
    if (const auto* a = std::get_if<Array>(&v_)) return a->size();
    if (const auto* o = std::get_if<ObjectPtr>(&v_)) return (*o)->size();
    if (const auto* s = std::get_if<std::string>(&v_)) return s->size();
    return 0;
  }

 private:
  // Objects are held through a pointer: the map's value type is Value
  // itself, incomplete here.
  using ObjectPtr = std::shared_ptr<const Object>;
  std::variant<std::nullptr_t, bool, double, std::string, Array, ObjectPtr, Numbers> v_ = nullptr;
  mutable std::shared_ptr<Array> expanded_;
};

namespace detail {

class Parser {
 public:
  Parser(const char* b, const char* e) : p_(b), b_(b), e_(e) {}

  Value document() {
    Value v = value(0);
    ws();
    if (p_ != e_) fail("trailing characters");
    return v;
  }

 private:
  const char* p_;
  const char* b_;
  const char* e_;

  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error("JSON parse error at offset " + std::to_string(p_ - b_) + ": " +
                             what);
  }
  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
  }
  bool eat(char c) {
    ws();
    if (p_ < e_ && *p_ == c) {
      ++p_;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail(std::string("expected '") + c + "'");
  }
  void literal(const char* word) {
    const size_t n = std::strlen(word);
    if (static_cast<size_t>(e_ - p_) < n || std::strncmp(p_, word, n) != 0)
      fail(std::string("expected ") + word);
    p_ += n;
  }

  Value value(int depth) {
    if (depth > 512) fail("nesting too deep");
    ws();
    if (p_ >= e_) fail("unexpected end of input");
    switch (*p_) {
      case '{': return object(depth);
      case '[': return array(depth);
      case '"': return Value(string());
      case 't': literal("true"); return Value(true);
      case 'f': literal("false"); return Value(false);
      case 'n': literal("null"); return Value(nullptr);
      default: return Value(number());
    }
  }

  Value object(int depth) {
    ++p_;  // '{'
    Value::Object o;
    if (eat('}')) return Value(std::move(o));
    do {
      ws();
      if (p_ >= e_ || *p_ != '"') fail("expected object key");
      std::string key = string();
      expect(':');
      o.insert_or_assign(std::move(key), value(depth + 1));
    } while (eat(','));
    expect('}');
    return Value(std::move(o));
  }

  // Arrays of plain numbers are read straight into a packed vector; the
  // first non-number element switches to a general array.
  Value array(int depth) {
    ++p_;  // '['
    if (eat(']')) return Value(Value::Array{});
    Value::Numbers nums;
    for (;;) {
      ws();
      if (p_ < e_ && (*p_ == '-' || (*p_ >= '0' && *p_ <= '9'))) {
        nums.push_back(number());
        if (eat(',')) continue;
        expect(']');
        return Value::numbers(std::move(nums));
      }
      break;
    }
    Value::Array a;
    a.reserve(nums.size() + 1);
    for (double d : nums) a.emplace_back(d);
    do {
      a.push_back(value(depth + 1));
    } while (eat(','));
    expect(']');
    return Value(std::move(a));
  }

  double number() {
    const char* s = p_;
    if (p_ < e_ && *p_ == '-') ++p_;
    if (p_ >= e_ || !(*p_ >= '0' && *p_ <= '9')) fail("invalid number");
    while (p_ < e_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
                       *p_ == '+' || *p_ == '-'))
      ++p_;
    std::string tok(s, p_);
    char* end = nullptr;
    const double d = std::strtod(tok.c_str(), &end);
    if (end != tok.c_str() + tok.size()) fail("invalid number '" + tok + "'");
    return d;
  }

  static void put_utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) {
      out += static_cast<char>(cp);
    } else if (cp < 0x800) {
      out += static_cast<char>(0xC0 | (cp >> 6));
      out += static_cast<char>(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += static_cast<char>(0xE0 | (cp >> 12));
      out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
      out += static_cast<char>(0x80 | (cp & 0x3F));
    } else {
      out += static_cast<char>(0xF0 | (cp >> 18));
      out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
      out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
      out += static_cast<char>(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (e_ - p_ < 4) fail("truncated \\u escape");
    unsigned v = 0;
    for (int n = 0; n < 4; ++n) {
      const char c = *p_++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= static_cast<unsigned>(c - '0');
      else if (c >= 'a' && c <= 'f') v |= static_cast<unsigned>(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= static_cast<unsigned>(c - 'A' + 10);
      else fail("invalid \\u escape");
    }
    return v;
  }
  std::string string() {
    ++p_;  // '"'
    std::string out;
    for (;;) {
      if (p_ >= e_) fail("unterminated string");
      const char c = *p_++;
      if (c == '"') return out;
      if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= e_) fail("unterminated escape");
      const char x = *p_++;
      switch (x) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
            if (e_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u') fail("unpaired surrogate");
            p_ += 2;
            const unsigned lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) fail("invalid low surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(out, cp);
          break;
        }
        default: fail("invalid escape");
      }
    }
  }
};

}  // namespace detail

inline Value parse_text(const std::string& text) {
  detail::Parser p(text.data(), text.data() + text.size());
  return p.document();
}

inline Value parse_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("Failed to open JSON file: " + path);
  std::ostringstream buf;
  buf << in.rdbuf();
  return parse_text(buf.str());
}

inline std::vector<real> to_real_vector(const Value& value) {
  std::vector<real> out;
  if (value.is_numbers()) {
    const auto& n = value.as_numbers();
    out.reserve(n.size());
    for (double d : n) out.push_back(static_cast<real>(d));
    return out;
  }
  const auto& a = value.as_array();
  out.reserve(a.size());
  for (const auto& e : a) out.push_back(static_cast<real>(e.as_number()));
  return out;
}

inline std::unordered_map<std::string, std::string> to_string_map(const Value& value) {
  std::unordered_map<std::string, std::string> out;
  for (const auto& [k, e] : value.as_object()) {
    if (e.is_string()) out.emplace(k, e.as_string());
    else if (e.is_number()) out.emplace(k, std::to_string(e.as_number()));
    else if (e.is_bool()) out.emplace(k, e.as_bool() ? "true" : "false");
    else if (e.is_null()) out.emplace(k, "null");
  }
  return out;
}

}  // namespace gwm::ingest::json
