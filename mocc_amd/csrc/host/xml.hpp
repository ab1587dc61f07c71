// Minimal XML DOM for mocc input decks.
//
// mocc reads its input with pugixml (src/input_proc.cpp:51-162); decks have
// several top-level elements and no single root, e.g. examples/3x3.xml. This
// reader keeps exactly what those decks use: elements, attributes, character
// data, comments, CDATA and the five predefined entities. Attribute
// conversions follow pugixml's documented rules (as_bool: true when the first
// character is one of 1tTyY; as_int/as_double: strtol/strtod with a default
// on empty input).
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace mocc_amd {

struct XmlNode {
    std::string name;  // empty for the document node
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XmlNode>> children;
    std::string text;  // first character-data child (pugixml child_value())
    XmlNode *parent = nullptr;

    // First child element with the given name (nullptr when absent).
    const XmlNode *child(const std::string &n) const;
    std::vector<const XmlNode *> children_named(const std::string &n) const;
    bool has_attr(const std::string &n) const;
    const std::string *attr(const std::string &n) const;
    std::string attr_str(const std::string &n, const std::string &def = "") const;
    int attr_int(const std::string &n, int def) const;
    double attr_double(const std::string &n, double def) const;
    bool attr_bool(const std::string &n, bool def) const;
};

// Parse a whole document; throws std::runtime_error with the byte offset on
// malformed input.
std::unique_ptr<XmlNode> xml_parse_string(const std::string &doc);
std::unique_ptr<XmlNode> xml_parse_file(const std::string &path);

}  // namespace mocc_amd
