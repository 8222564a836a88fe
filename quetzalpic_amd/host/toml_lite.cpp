#include "toml_lite.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace toml_lite {

bool Table::get(const std::string &k, double &out) const {
    auto it = values.find(k);
    if (it == values.end()) return false;
    if (it->second.kind == Value::FLOAT) out = it->second.f;
    else if (it->second.kind == Value::INT) out = (double)it->second.i;
    else return false;
    return true;
}
bool Table::get(const std::string &k, int &out) const {
    auto it = values.find(k);
    if (it == values.end() || it->second.kind != Value::INT) return false;
    out = (int)it->second.i;
    return true;
}
bool Table::get(const std::string &k, bool &out) const {
    auto it = values.find(k);
    if (it == values.end() || it->second.kind != Value::BOOL) return false;
    out = it->second.b;
    return true;
}
bool Table::get(const std::string &k, std::string &out) const {
    auto it = values.find(k);
    if (it == values.end() || it->second.kind != Value::STRING) return false;
    out = it->second.s;
    return true;
}

static std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

// Remove a trailing comment that is not inside a string.
static std::string strip_comment(const std::string &line) {
    char q = 0;
    for (size_t i = 0; i < line.size(); ++i) {
        char c = line[i];
        if (q) {
            if (c == '\\' && q == '"') ++i;
            else if (c == q) q = 0;
        } else if (c == '"' || c == '\'') {
            q = c;
        } else if (c == '#') {
            return line.substr(0, i);
        }
    }
    return line;
}

static Value parse_value(const std::string &raw, int lineno) {
    std::string t = trim(raw);
    Value v;
    v.text = t;
    auto fail = [&](const std::string &why) {
        throw parse_error("line " + std::to_string(lineno) + ": " + why + ": " + t);
    };
    if (t.empty()) fail("missing value");
    if (t[0] == '"' || t[0] == '\'') {
        const char q = t[0];
        if (t.size() < 2 || t.back() != q) fail("unterminated string");
        std::string s;
        for (size_t i = 1; i + 1 < t.size(); ++i) {
            char c = t[i];
            if (q == '"' && c == '\\' && i + 2 < t.size()) {
                char e = t[++i];
                s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
            } else {
                s += c;
            }
        }
        v.kind = Value::STRING;
        v.s = s;
        return v;
    }
    if (t == "true" || t == "false") {
        v.kind = Value::BOOL;
        v.b = (t == "true");
        return v;
    }
    std::string num;
    for (char c : t)
        if (c != '_') num += c;
    bool is_float = num.find_first_of(".eE") != std::string::npos || num == "inf" || num == "+inf" ||
                    num == "-inf" || num == "nan";
    errno = 0;
    char *end = nullptr;
    if (is_float) {
        double d = std::strtod(num.c_str(), &end);
        if (end == num.c_str() || *end != 0) fail("bad number");
        v.kind = Value::FLOAT;
        v.f = d;
    } else {
        long long i = std::strtoll(num.c_str(), &end, 10);
        if (end == num.c_str() || *end != 0 || errno) fail("bad value");
        v.kind = Value::INT;
        v.i = i;
    }
    return v;
}

TablePtr parse_string(const std::string &text) {
    auto root = std::make_shared<Table>();
    TablePtr cur = root;
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        std::string s = trim(strip_comment(line));
        if (s.empty()) continue;
        if (s.compare(0, 2, "[[") == 0) {
            if (s.size() < 4 || s.compare(s.size() - 2, 2, "]]") != 0)
                throw parse_error("line " + std::to_string(lineno) + ": bad table array header");
            std::string name = trim(s.substr(2, s.size() - 4));
            if (root->values.count(name) || root->tables.count(name))
                throw parse_error("line " + std::to_string(lineno) + ": [[" + name + "]] redefines a key");
            cur = std::make_shared<Table>();
            root->table_arrays[name].push_back(cur);
            continue;
        }
        if (s[0] == '[') {
            if (s.back() != ']') throw parse_error("line " + std::to_string(lineno) + ": bad table header");
            const std::string name = trim(s.substr(1, s.size() - 2));
            // [a.b]: table b inside table a, or inside the last element of [[a]]
            TablePtr parent = root;
            std::string last = name;
            for (size_t dot; (dot = last.find('.')) != std::string::npos;) {
                const std::string part = trim(last.substr(0, dot));
                last = trim(last.substr(dot + 1));
                if (parent->is_table_array(part)) {
                    parent = parent->table_arrays[part].back();
                } else if (parent->is_table(part)) {
                    parent = parent->tables[part];
                } else if (!parent->contains(part) && !part.empty()) {
                    auto t = std::make_shared<Table>();
                    parent->tables[part] = t;
                    parent = t;
                } else {
                    throw parse_error("line " + std::to_string(lineno) + ": bad table header " + name);
                }
            }
            if (last.empty()) throw parse_error("line " + std::to_string(lineno) + ": bad table header " + name);
            if (parent->contains(last))
                throw parse_error("line " + std::to_string(lineno) + ": duplicate table " + name);
            cur = std::make_shared<Table>();
            parent->tables[last] = cur;
            continue;
        }
        size_t eq = s.find('=');
        if (eq == std::string::npos) throw parse_error("line " + std::to_string(lineno) + ": expected key = value");
        std::string key = trim(s.substr(0, eq));
        if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'') && key.back() == key[0])
            key = key.substr(1, key.size() - 2);
        if (key.empty()) throw parse_error("line " + std::to_string(lineno) + ": empty key");
        if (cur->values.count(key)) throw parse_error("line " + std::to_string(lineno) + ": duplicate key " + key);
        cur->values[key] = parse_value(s.substr(eq + 1), lineno);
    }
    return root;
}

TablePtr parse_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw parse_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_string(ss.str());
}

}  // namespace toml_lite
