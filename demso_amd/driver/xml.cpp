// xml.cpp -- recursive-descent reader behind xml.hpp.
#include "xml.hpp"

#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>

namespace dsml {

const XmlNode* XmlNode::child(const std::string& n) const {
  for (const auto& c : children)
    if (c->name == n) return c.get();
  return nullptr;
}

const std::string* XmlNode::attr(const std::string& n) const {
  for (const auto& a : attrs)
    if (a.first == n) return &a.second;
  return nullptr;
}

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

bool is_name_char(char c) {
  const unsigned char u = (unsigned char)c;
  return std::isalnum(u) || c == '_' || c == '-' || c == '.' || c == ':' || u >= 0x80;
}

class Reader {
 public:
  explicit Reader(const std::string& s) : s_(s) {}

  std::unique_ptr<XmlNode> document() {
    skip_misc();
    if (!peek("<") || peek("</")) fail("expected the root element");
    auto root = element();
    skip_misc();
    if (p_ != s_.size()) fail("content after the root element");
    return root;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;
  int line_ = 1;

  [[noreturn]] void fail(const std::string& what) const {
    throw XmlError("XML error at line " + std::to_string(line_) + ": " + what);
  }

  bool peek(const char* lit) const { return s_.compare(p_, std::strlen(lit), lit) == 0; }

  void advance(size_t k) {
    for (size_t i = 0; i < k && p_ < s_.size(); ++i, ++p_)
      if (s_[p_] == '\n') ++line_;
  }

  void skip_space() {
    while (p_ < s_.size() && is_space(s_[p_])) advance(1);
  }

  // skip up to and including `end`
  void skip_past(const char* end, const char* what) {
    const size_t q = s_.find(end, p_);
    if (q == std::string::npos) fail(std::string("unterminated ") + what);
    advance(q + std::strlen(end) - p_);
  }

  // declaration, processing instructions, comments, DOCTYPE, whitespace
  void skip_misc() {
    for (;;) {
      skip_space();
      if (peek("<?")) skip_past("?>", "processing instruction");
      else if (peek("<!--")) skip_past("-->", "comment");
      else if (peek("<!DOCTYPE")) skip_past(">", "DOCTYPE");
      else return;
    }
  }

  std::string name() {
    const size_t b = p_;
    while (p_ < s_.size() && is_name_char(s_[p_])) advance(1);
    if (p_ == b) fail("expected a name");
    return s_.substr(b, p_ - b);
  }

  static void put_utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }

  // decode an entity reference starting at '&'
  void entity(std::string& out) {
    const size_t semi = s_.find(';', p_);
    if (semi == std::string::npos || semi - p_ > 12) fail("unterminated entity reference");
    const std::string ref = s_.substr(p_ + 1, semi - p_ - 1);
    if (ref == "lt") out += '<';
    else if (ref == "gt") out += '>';
    else if (ref == "amp") out += '&';
    else if (ref == "quot") out += '"';
    else if (ref == "apos") out += '\'';
    else if (ref.size() > 1 && ref[0] == '#') {
      const bool hex = ref[1] == 'x' || ref[1] == 'X';
      const std::string digits = ref.substr(hex ? 2 : 1);
      if (digits.empty()) fail("empty character reference");
      uint32_t cp = 0;
      for (char c : digits) {
        int v;
        if (c >= '0' && c <= '9') v = c - '0';
        else if (hex && c >= 'a' && c <= 'f') v = c - 'a' + 10;
        else if (hex && c >= 'A' && c <= 'F') v = c - 'A' + 10;
        else fail("bad character reference &" + ref + ";");
        cp = cp * (hex ? 16 : 10) + (uint32_t)v;
        if (cp > 0x10FFFF) fail("character reference out of range");
      }
      put_utf8(out, cp);
    } else {
      fail("unknown entity &" + ref + ";");
    }
    advance(semi + 1 - p_);
  }

  std::string attr_value() {
    if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) fail("expected a quoted attribute value");
    const char q = s_[p_];
    advance(1);
    std::string v;
    while (p_ < s_.size() && s_[p_] != q) {
      if (s_[p_] == '<') fail("'<' in attribute value");
      if (s_[p_] == '&') entity(v);
      else {
        v += s_[p_];
        advance(1);
      }
    }
    if (p_ >= s_.size()) fail("unterminated attribute value");
    advance(1);
    return v;
  }

  static bool blank(const std::string& t) {
    for (char c : t)
      if (!is_space(c)) return false;
    return true;
  }

  // at '<' of a start tag
  std::unique_ptr<XmlNode> element() {
    auto node = std::make_unique<XmlNode>();
    node->line = line_;
    advance(1);
    node->name = name();
    for (;;) {
      const bool had_space = p_ < s_.size() && is_space(s_[p_]);
      skip_space();
      if (p_ >= s_.size()) fail("unterminated start tag <" + node->name + ">");
      if (peek("/>")) {
        advance(2);
        return node;
      }
      if (s_[p_] == '>') {
        advance(1);
        break;
      }
      if (!had_space) fail("expected whitespace before an attribute of <" + node->name + ">");
      std::string an = name();
      skip_space();
      if (p_ >= s_.size() || s_[p_] != '=') fail("expected '=' after attribute " + an);
      advance(1);
      skip_space();
      if (node->attr(an)) fail("duplicate attribute " + an);
      node->attrs.emplace_back(std::move(an), attr_value());
    }
    // content
    std::string run;  // character data since the last child element
    auto flush = [&]() {
      if (!blank(run)) {
        node->text += run;
        auto d = std::make_unique<XmlNode>();
        d->text = run;
        d->line = line_;
        node->children.push_back(std::move(d));
      }
      run.clear();
    };
    for (;;) {
      if (p_ >= s_.size()) fail("missing </" + node->name + ">");
      if (peek("</")) {
        flush();
        advance(2);
        const std::string close = name();
        if (close != node->name) fail("</" + close + "> closes <" + node->name + ">");
        skip_space();
        if (p_ >= s_.size() || s_[p_] != '>') fail("expected '>' in </" + close + ">");
        advance(1);
        return node;
      }
      if (peek("<!--")) {
        skip_past("-->", "comment");
      } else if (peek("<![CDATA[")) {
        advance(9);
        const size_t q = s_.find("]]>", p_);
        if (q == std::string::npos) fail("unterminated CDATA section");
        run += s_.substr(p_, q - p_);
        advance(q + 3 - p_);
      } else if (peek("<?")) {
        skip_past("?>", "processing instruction");
      } else if (s_[p_] == '<') {
        flush();
        node->children.push_back(element());
      } else if (s_[p_] == '&') {
        entity(run);
      } else {
        run += s_[p_];
        advance(1);
      }
    }
  }
};

}  // namespace

std::unique_ptr<XmlNode> parse_xml(const std::string& doc) { return Reader(doc).document(); }

std::unique_ptr<XmlNode> parse_xml_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw XmlError("cannot open " + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return parse_xml(ss.str());
}

}  // namespace dsml
