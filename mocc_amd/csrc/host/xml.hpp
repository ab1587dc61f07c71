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

    XmlNode *child_mut(const std::string &n);
    std::string *attr_mut(const std::string &n);
    bool has_text = false;  // the element holds character data (pugixml's xml_text is non-empty)
};

// Parse a whole document; throws std::runtime_error with the byte offset on
// malformed input.
std::unique_ptr<XmlNode> xml_parse_string(const std::string &doc);
std::unique_ptr<XmlNode> xml_parse_file(const std::string &path);

// The document as XML text: elements, attributes and character data in
// document order (comments and processing instructions are not kept), two
// spaces of indentation per level, as pugixml's save() lays a document out.
std::string xml_serialize(const XmlNode &doc);

// A command-line amendment "path/to/name=value" (InputProcessor,
// src/input_proc.cpp:127-146, 217-272): every path component but the last
// names a child element, followed from the document node; the last names
// an attribute or a child element of that element (exactly one of the two
// must exist), whose value or character data becomes `value`. Throws
// std::runtime_error with mocc's messages on a malformed amendment or a
// missing target.
void xml_apply_amendment(XmlNode &doc, const std::string &amendment);

}  // namespace mocc_amd
