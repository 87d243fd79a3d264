// xml.hpp -- a small XML reader for .dsml scene files.
//
// Reads a whole document into a tree of elements.  Supported: the XML
// declaration and processing instructions (skipped), comments (skipped),
// DOCTYPE (skipped), CDATA, elements with single- or double-quoted
// attributes, self-closing elements, character data and the five predefined
// entities plus numeric character references.  Character data that is only
// whitespace is dropped; an element's text() is the concatenation of its
// direct character data (what a .dsml reader reads as the node's value).
// Malformed input raises XmlError with the line number.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dsml {

struct XmlError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct XmlNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::string text;  // direct character data, entities decoded
  // Children in document order.  A non-whitespace run of character data
  // between child elements is kept as a child with an empty name.
  std::vector<std::unique_ptr<XmlNode>> children;
  int line = 0;

  const XmlNode* child(const std::string& n) const;       // first child element named n, or null
  const std::string* attr(const std::string& n) const;    // attribute value, or null
};

// Parse a document held in memory; returns the root element.
std::unique_ptr<XmlNode> parse_xml(const std::string& doc);
// Read and parse a file (XmlError if it cannot be read).
std::unique_ptr<XmlNode> parse_xml_file(const std::string& path);

}  // namespace dsml
