#include "xml.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace mocc_amd {

const XmlNode *XmlNode::child(const std::string &n) const
{
    for (const auto &c : children) {
        if (c->name == n) return c.get();
    }
    return nullptr;
}

std::vector<const XmlNode *> XmlNode::children_named(const std::string &n) const
{
    std::vector<const XmlNode *> out;
    for (const auto &c : children) {
        if (c->name == n) out.push_back(c.get());
    }
    return out;
}

const std::string *XmlNode::attr(const std::string &n) const
{
    for (const auto &a : attrs) {
        if (a.first == n) return &a.second;
    }
    return nullptr;
}

bool XmlNode::has_attr(const std::string &n) const { return attr(n) != nullptr; }

std::string XmlNode::attr_str(const std::string &n, const std::string &def) const
{
    const std::string *v = attr(n);
    return v ? *v : def;
}

int XmlNode::attr_int(const std::string &n, int def) const
{
    const std::string *v = attr(n);
    if (!v || v->empty()) return def;
    return (int)std::strtol(v->c_str(), nullptr, 10);
}

double XmlNode::attr_double(const std::string &n, double def) const
{
    const std::string *v = attr(n);
    if (!v || v->empty()) return def;
    return std::strtod(v->c_str(), nullptr);
}

bool XmlNode::attr_bool(const std::string &n, bool def) const
{
    const std::string *v = attr(n);
    if (!v || v->empty()) return def;
    char c = (*v)[0];
    return c == '1' || c == 't' || c == 'T' || c == 'y' || c == 'Y';
}

XmlNode *XmlNode::child_mut(const std::string &n)
{
    for (auto &c : children)
        if (c->name == n) return c.get();
    return nullptr;
}

std::string *XmlNode::attr_mut(const std::string &n)
{
    for (auto &a : attrs)
        if (a.first == n) return &a.second;
    return nullptr;
}

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;

    [[noreturn]] void fail(const char *what) const
    {
        std::ostringstream m;
        m << "XML parse error at offset " << i << ": " << what;
        throw std::runtime_error(m.str());
    }

    bool starts(const char *p) const { return s.compare(i, std::char_traits<char>::length(p), p) == 0; }

    void skip_ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }

    static std::string decode(const std::string &raw)
    {
        std::string out;
        out.reserve(raw.size());
        for (size_t k = 0; k < raw.size(); k++) {
            if (raw[k] != '&') {
                out += raw[k];
                continue;
            }
            size_t e = raw.find(';', k);
            if (e == std::string::npos) {
                out += raw[k];
                continue;
            }
            std::string ent = raw.substr(k + 1, e - k - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                long cp = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X'))
                              ? std::strtol(ent.c_str() + 2, nullptr, 16)
                              : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (cp < 128) out += (char)cp;
            } else {
                out += raw.substr(k, e - k + 1);
            }
            k = e;
        }
        return out;
    }

    std::string name()
    {
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' ||
                                s[i] == '-' || s[i] == ':' || s[i] == '.'))
            i++;
        if (b == i) fail("expected a name");
        return s.substr(b, i - b);
    }

    // Skip comments, processing instructions and doctype declarations.
    bool skip_misc()
    {
        if (starts("<!--")) {
            size_t e = s.find("-->", i + 4);
            if (e == std::string::npos) fail("unterminated comment");
            i = e + 3;
            return true;
        }
        if (starts("<?")) {
            size_t e = s.find("?>", i + 2);
            if (e == std::string::npos) fail("unterminated processing instruction");
            i = e + 2;
            return true;
        }
        if (starts("<!DOCTYPE")) {
            size_t e = s.find('>', i);
            if (e == std::string::npos) fail("unterminated doctype");
            i = e + 1;
            return true;
        }
        return false;
    }

    void add_text(XmlNode &n, const std::string &t, bool &have_text)
    {
        // pugixml drops whitespace-only character data by default.
        bool all_ws = true;
        for (char c : t) {
            if (!std::isspace((unsigned char)c)) {
                all_ws = false;
                break;
            }
        }
        if (all_ws || have_text) return;
        n.text = t;
        n.has_text = true;
        have_text = true;
    }

    void content(XmlNode &node, bool is_doc)
    {
        bool have_text = false;
        while (true) {
            if (i >= s.size()) {
                if (is_doc) return;
                fail("unexpected end of input");
            }
            if (s[i] != '<') {
                size_t e = s.find('<', i);
                if (e == std::string::npos) e = s.size();
                add_text(node, decode(s.substr(i, e - i)), have_text);
                i = e;
                continue;
            }
            if (skip_misc()) continue;
            if (starts("<![CDATA[")) {
                size_t e = s.find("]]>", i + 9);
                if (e == std::string::npos) fail("unterminated CDATA");
                add_text(node, s.substr(i + 9, e - i - 9), have_text);
                i = e + 3;
                continue;
            }
            if (starts("</")) {
                if (is_doc) fail("unmatched closing tag");
                i += 2;
                std::string n = name();
                if (n != node.name) fail("mismatched closing tag");
                skip_ws();
                if (i >= s.size() || s[i] != '>') fail("expected '>'");
                i++;
                return;
            }
            // element
            i++;
            auto el = std::make_unique<XmlNode>();
            el->parent = &node;
            el->name = name();
            while (true) {
                skip_ws();
                if (i >= s.size()) fail("unexpected end in tag");
                if (s[i] == '/') {
                    if (i + 1 >= s.size() || s[i + 1] != '>') fail("expected '/>'");
                    i += 2;
                    node.children.push_back(std::move(el));
                    break;
                }
                if (s[i] == '>') {
                    i++;
                    content(*el, false);
                    node.children.push_back(std::move(el));
                    break;
                }
                std::string an = name();
                skip_ws();
                if (i >= s.size() || s[i] != '=') fail("expected '='");
                i++;
                skip_ws();
                if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) fail("expected quote");
                char q = s[i++];
                size_t e = s.find(q, i);
                if (e == std::string::npos) fail("unterminated attribute");
                el->attrs.emplace_back(an, decode(s.substr(i, e - i)));
                i = e + 1;
            }
        }
    }
};

}  // namespace

std::unique_ptr<XmlNode> xml_parse_string(const std::string &doc)
{
    auto root = std::make_unique<XmlNode>();
    Parser p{doc};
    p.content(*root, true);
    return root;
}

std::unique_ptr<XmlNode> xml_parse_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open XML file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return xml_parse_string(ss.str());
}

namespace {
std::string escape(const std::string &v, bool attr)
{
    std::string o;
    o.reserve(v.size());
    for (char c : v) {
        if (c == '&') o += "&amp;";
        else if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else if (attr && c == '"') o += "&quot;";
        else o += c;
    }
    return o;
}

void serialize(const XmlNode &n, int depth, std::string &o)
{
    const std::string ind(2 * depth, ' ');
    o += ind + "<" + n.name;
    for (const auto &a : n.attrs) o += " " + a.first + "=\"" + escape(a.second, true) + "\"";
    if (n.children.empty() && !n.has_text) {
        o += " />\n";
        return;
    }
    o += ">";
    if (n.children.empty()) {
        o += escape(n.text, false) + "</" + n.name + ">\n";
        return;
    }
    o += "\n";
    if (n.has_text) o += ind + "  " + escape(n.text, false) + "\n";
    for (const auto &c : n.children) serialize(*c, depth + 1, o);
    o += ind + "</" + n.name + ">\n";
}
}  // namespace

std::string xml_serialize(const XmlNode &doc)
{
    std::string o = "<?xml version=\"1.0\"?>\n";
    for (const auto &c : doc.children) serialize(*c, 0, o);
    return o;
}

void xml_apply_amendment(XmlNode &doc, const std::string &amendment)
{
    const size_t loc = amendment.find('=');
    if (loc == std::string::npos)
        throw std::runtime_error("Malformed replacement in command line. Proper syntax:\n"
                                 "-a path/to/attribute:new_value");
    if (amendment.rfind('=') != loc) throw std::runtime_error("Malformed replacement in command line");
    std::string path = amendment.substr(0, loc);
    const std::string value = amendment.substr(loc + 1);
    std::vector<std::string> comp;
    while (true) {
        const size_t pos = path.find('/');
        comp.push_back(path.substr(0, pos));
        if (pos == std::string::npos) break;
        path = path.substr(pos + 1);
    }
    XmlNode *node = &doc;
    for (size_t k = 0; k + 1 < comp.size(); k++) {
        node = node->child_mut(comp[k]);
        if (!node) throw std::runtime_error("Failed to apply command-line amendment");
    }
    const std::string &last = comp.back();
    std::string *a = node->attr_mut(last);
    XmlNode *c = node->child_mut(last);
    if ((a != nullptr) == (c != nullptr))
        throw std::runtime_error("Could not find the requested attribute or child to modify, or could not "
                                 "disambiguate.");
    if (a) {
        *a = value;
    } else {
        if (!c->has_text) throw std::runtime_error("No text data found at requested location.");
        c->text = value;
    }
}

}  // namespace mocc_amd
