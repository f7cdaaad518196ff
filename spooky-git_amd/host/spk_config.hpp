// spk_config.hpp -- reader for the configuration files SPOOKY problems ship
// (spooky.cfg): the subset of the libconfig grammar those files use.
//
//   file    := setting*
//   setting := name ('=' | ':') value (';' | ',')?
//   value   := scalar | '[' scalars ']' | '(' values ')' | '{' setting* '}'
//   scalar  := bool | int | float | string ("..." "..." concatenates)
//   comments: '#' and '//' to end of line, '/* ... */'
//
// Lookups take libconfig paths: "physics.reynolds", "physics.boxsize.[0]".
// A lookup of a missing path, or of a value of another kind, returns false
// and leaves the output untouched -- the reference's config_lookup_* pattern
// (parameters.cu:77-505) keeps its defaults exactly that way.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace spk {

struct CfgValue {
    enum Kind { BOOL, INT, FLOAT, STRING, ARRAY, LIST, GROUP } kind = GROUP;
    bool b = false;
    long long i = 0;
    double f = 0.0;
    std::string s;
    std::vector<std::shared_ptr<CfgValue>> elems;               // ARRAY, LIST
    std::vector<std::pair<std::string, std::shared_ptr<CfgValue>>> members;  // GROUP, in file order
    int line = 0;

    const CfgValue *member(const std::string &name) const;
};

class Config {
public:
    // Parses text; on a syntax error returns false and sets error() to
    // "line N: message".
    bool parse(const std::string &text);
    bool read_file(const std::string &path);
    const std::string &error() const { return err_; }

    const CfgValue *lookup(const std::string &path) const;
    // libconfig without auto-conversion, as the reference configures it
    // (parameters.cu:59): an int setting does not satisfy a float lookup nor a
    // float setting an int lookup, so "reynolds = 100;" keeps the default.
    bool lookup_bool(const std::string &path, int *out) const;
    bool lookup_int(const std::string &path, int *out) const;
    bool lookup_float(const std::string &path, double *out) const;
    bool lookup_string(const std::string &path, std::string *out) const;
    // strings of an array or list setting; false when absent
    bool lookup_strings(const std::string &path, std::vector<std::string> *out) const;

private:
    std::shared_ptr<CfgValue> root_;
    std::string err_;
};

}  // namespace spk
