// json.h — a small JSON reader for the glTF loader: parses a document into a
// tree of values (RFC 8259: objects, arrays, strings with escapes, numbers,
// true/false/null).  Objects keep their members in document order.
#ifndef CG_HOST_JSON_H_
#define CG_HOST_JSON_H_

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace crystal::graphics::json {

struct Value {
  enum Kind { kNull, kBool, kNumber, kString, kArray, kObject } kind = kNull;
  bool b = false;
  double num = 0.0;
  std::string str;  // a string's value; a number's text as written
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  const Value* get(std::string_view key) const {
    if (kind != kObject) return nullptr;
    for (const auto& m : obj)
      if (m.first == key) return &m.second;
    return nullptr;
  }
  const Value* at(size_t i) const { return kind == kArray && i < arr.size() ? &arr[i] : nullptr; }
  size_t size() const { return kind == kArray ? arr.size() : 0; }
  bool is_num() const { return kind == kNumber; }
  double number(double dflt) const { return kind == kNumber ? num : dflt; }
};

class Parser {
 public:
  explicit Parser(std::string_view text) : s_(text) {}
  // Parses the whole text; on failure returns false and sets error().
  bool parse(Value& out) {
    skip();
    if (!value(out, 0)) return false;
    skip();
    if (p_ != s_.size()) return fail("trailing characters");
    return true;
  }
  const std::string& error() const { return err_; }

 private:
  std::string_view s_;
  size_t p_ = 0;
  std::string err_;

  bool fail(const char* what) {
    if (err_.empty()) err_ = std::string("JSON: ") + what + " at offset " + std::to_string(p_);
    return false;
  }
  void skip() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
  }
  bool literal(std::string_view word) {
    if (s_.substr(p_, word.size()) != word) return fail("bad literal");
    p_ += word.size();
    return true;
  }
  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out.push_back(static_cast<char>(cp));
    } else if (cp < 0x800) {
      out.push_back(static_cast<char>(0xC0 | (cp >> 6)));
      out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
    } else if (cp < 0x10000) {
      out.push_back(static_cast<char>(0xE0 | (cp >> 12)));
      out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
      out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
    } else {
      out.push_back(static_cast<char>(0xF0 | (cp >> 18)));
      out.push_back(static_cast<char>(0x80 | ((cp >> 12) & 0x3F)));
      out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
      out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
    }
  }
  bool hex4(uint32_t& v) {
    if (p_ + 4 > s_.size()) return fail("short \\u escape");
    v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = s_[p_++];
      v <<= 4;
      if (c >= '0' && c <= '9')
        v |= static_cast<uint32_t>(c - '0');
      else if (c >= 'a' && c <= 'f')
        v |= static_cast<uint32_t>(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F')
        v |= static_cast<uint32_t>(c - 'A' + 10);
      else
        return fail("bad \\u escape");
    }
    return true;
  }
  bool string(std::string& out) {
    ++p_;  // opening quote
    while (p_ < s_.size()) {
      const char c = s_[p_++];
      if (c == '"') return true;
      if (c != '\\') {
        out.push_back(c);
        continue;
      }
      if (p_ >= s_.size()) break;
      const char e = s_[p_++];
      switch (e) {
        case '"': case '\\': case '/': out.push_back(e); break;
        case 'b': out.push_back('\b'); break;
        case 'f': out.push_back('\f'); break;
        case 'n': out.push_back('\n'); break;
        case 'r': out.push_back('\r'); break;
        case 't': out.push_back('\t'); break;
        case 'u': {
          uint32_t cp;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && p_ + 6 <= s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
            p_ += 2;
            uint32_t lo;
            if (!hex4(lo)) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: return fail("bad escape");
      }
    }
    return fail("unterminated string");
  }
  bool number(double& out, std::string& text) {
    const size_t b = p_;
    if (p_ < s_.size() && s_[p_] == '-') ++p_;
    while (p_ < s_.size() && ((s_[p_] >= '0' && s_[p_] <= '9') || s_[p_] == '.' || s_[p_] == 'e' || s_[p_] == 'E' ||
                              s_[p_] == '+' || s_[p_] == '-'))
      ++p_;
    const std::string tok(s_.substr(b, p_ - b));
    char* end = nullptr;
    out = std::strtod(tok.c_str(), &end);
    if (tok.empty() || end != tok.c_str() + tok.size()) return fail("bad number");
    text = tok;
    return true;
  }
  bool value(Value& v, int depth) {
    if (depth > 256) return fail("nesting too deep");
    skip();
    if (p_ >= s_.size()) return fail("unexpected end");
    const char c = s_[p_];
    if (c == '{') {
      v.kind = Value::kObject;
      ++p_;
      skip();
      if (p_ < s_.size() && s_[p_] == '}') {
        ++p_;
        return true;
      }
      for (;;) {
        skip();
        if (p_ >= s_.size() || s_[p_] != '"') return fail("expected member name");
        std::string key;
        if (!string(key)) return false;
        skip();
        if (p_ >= s_.size() || s_[p_] != ':') return fail("expected ':'");
        ++p_;
        v.obj.emplace_back(std::move(key), Value{});
        if (!value(v.obj.back().second, depth + 1)) return false;
        skip();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == '}') {
          ++p_;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Value::kArray;
      ++p_;
      skip();
      if (p_ < s_.size() && s_[p_] == ']') {
        ++p_;
        return true;
      }
      for (;;) {
        v.arr.emplace_back();
        if (!value(v.arr.back(), depth + 1)) return false;
        skip();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == ']') {
          ++p_;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = Value::kString;
      return string(v.str);
    }
    if (c == 't') {
      v.kind = Value::kBool;
      v.b = true;
      return literal("true");
    }
    if (c == 'f') {
      v.kind = Value::kBool;
      return literal("false");
    }
    if (c == 'n') return literal("null");
    v.kind = Value::kNumber;
    return number(v.num, v.str);
  }
};

}  // namespace crystal::graphics::json

#endif
