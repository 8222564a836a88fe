// toml_lite.h -- the subset of TOML the input decks use: comments, key = value
// with integers, floats, booleans and strings, [table], [[array of tables]] and
// dotted [a.b] headers (table b inside table a or the last element of [[a]]).
// Type rules follow what the reference's reader accepts: a double key takes an
// integer or a float, an integer key only an integer.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace toml_lite {

struct parse_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Value {
    enum Kind { INT, FLOAT, BOOL, STRING } kind = INT;
    long long i = 0;
    double f = 0;
    bool b = false;
    std::string s;
    std::string text;  // as written, for echoing
};

class Table;
using TablePtr = std::shared_ptr<Table>;

class Table {
public:
    std::map<std::string, Value> values;
    std::map<std::string, TablePtr> tables;
    std::map<std::string, std::vector<TablePtr>> table_arrays;

    bool contains(const std::string &k) const {
        return values.count(k) || tables.count(k) || table_arrays.count(k);
    }
    bool is_table(const std::string &k) const { return tables.count(k) > 0; }
    bool is_table_array(const std::string &k) const { return table_arrays.count(k) > 0; }
    TablePtr get_table(const std::string &k) const {
        auto it = tables.find(k);
        return it == tables.end() ? nullptr : it->second;
    }
    const std::vector<TablePtr> *get_table_array(const std::string &k) const {
        auto it = table_arrays.find(k);
        return it == table_arrays.end() ? nullptr : &it->second;
    }
    // Typed access: returns false when missing or of the wrong type.
    bool get(const std::string &k, double &out) const;
    bool get(const std::string &k, int &out) const;
    bool get(const std::string &k, bool &out) const;
    bool get(const std::string &k, std::string &out) const;
};

TablePtr parse_string(const std::string &text);
TablePtr parse_file(const std::string &path);

}  // namespace toml_lite
