// spk_config.cpp -- libconfig-subset reader (grammar in spk_config.hpp).
#include "spk_config.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace spk {

const CfgValue *CfgValue::member(const std::string &name) const {
    if (kind != GROUP) return nullptr;
    for (const auto &m : members)
        if (m.first == name) return m.second.get();
    return nullptr;
}

namespace {

struct Lexer {
    const std::string &t;
    size_t p = 0;
    int line = 1;
    std::string err;
    explicit Lexer(const std::string &text) : t(text) {}

    bool fail(const std::string &m) {
        if (err.empty()) err = "line " + std::to_string(line) + ": " + m;
        return false;
    }
    // skip blanks and comments
    void skip() {
        for (;;) {
            while (p < t.size() && std::isspace((unsigned char)t[p])) {
                if (t[p] == '\n') line++;
                p++;
            }
            if (p < t.size() && t[p] == '#') {
                while (p < t.size() && t[p] != '\n') p++;
            } else if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '/') {
                while (p < t.size() && t[p] != '\n') p++;
            } else if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '*') {
                p += 2;
                while (p + 1 < t.size() && !(t[p] == '*' && t[p + 1] == '/')) {
                    if (t[p] == '\n') line++;
                    p++;
                }
                p = p + 2 <= t.size() ? p + 2 : t.size();
            } else {
                return;
            }
        }
    }
    bool peek(char c) {
        skip();
        return p < t.size() && t[p] == c;
    }
    bool eat(char c) {
        if (peek(c)) {
            p++;
            return true;
        }
        return false;
    }
    bool name(std::string *out) {
        skip();
        size_t b = p;
        if (p < t.size() && (std::isalpha((unsigned char)t[p]) || t[p] == '*')) {
            p++;
            while (p < t.size() && (std::isalnum((unsigned char)t[p]) || t[p] == '_' || t[p] == '-' || t[p] == '*'))
                p++;
        }
        if (p == b) return fail("setting name expected");
        *out = t.substr(b, p - b);
        return true;
    }
    bool string(std::string *out) {
        out->clear();
        bool any = false;
        while (peek('"')) {
            any = true;
            p++;
            while (p < t.size() && t[p] != '"') {
                char c = t[p++];
                if (c == '\n') line++;
                if (c == '\\' && p < t.size()) {
                    char e = t[p++];
                    switch (e) {
                        case 'n': c = '\n'; break;
                        case 't': c = '\t'; break;
                        case 'r': c = '\r'; break;
                        case 'f': c = '\f'; break;
                        case '\\': c = '\\'; break;
                        case '"': c = '"'; break;
                        default: return fail(std::string("unknown escape \\") + e);
                    }
                }
                out->push_back(c);
            }
            if (p >= t.size()) return fail("unterminated string");
            p++;
        }
        return any ? true : fail("string expected");
    }
    // bool, int or float
    bool scalar(CfgValue *v) {
        skip();
        v->line = line;
        if (p < t.size() && t[p] == '"') {
            v->kind = CfgValue::STRING;
            return string(&v->s);
        }
        size_t b = p;
        while (p < t.size() && (std::isalnum((unsigned char)t[p]) || t[p] == '.' || t[p] == '+' || t[p] == '-'))
            p++;
        std::string w = t.substr(b, p - b);
        if (w.empty()) return fail("value expected");
        std::string lw;
        for (char c : w) lw.push_back((char)std::tolower((unsigned char)c));
        if (lw == "true" || lw == "false") {
            v->kind = CfgValue::BOOL;
            v->b = lw == "true";
            return true;
        }
        const bool hex = lw.size() > 2 && (lw.compare(0, 2, "0x") == 0 || lw.compare(0, 3, "-0x") == 0);
        const bool isfloat = !hex && (lw.find_first_of(".e") != std::string::npos);
        char *end = nullptr;
        if (isfloat) {
            v->kind = CfgValue::FLOAT;
            v->f = std::strtod(w.c_str(), &end);
        } else {
            std::string digits = w;
            while (!digits.empty() && (digits.back() == 'L' || digits.back() == 'l')) digits.pop_back();  // int64 suffix
            v->kind = CfgValue::INT;
            v->i = std::strtoll(digits.c_str(), &end, 0);
            if (end && *end == '\0') return true;
            return fail("bad number '" + w + "'");
        }
        if (!end || *end != '\0') return fail("bad number '" + w + "'");
        return true;
    }
    bool value(CfgValue *v) {
        skip();
        v->line = line;
        if (eat('[')) {
            v->kind = CfgValue::ARRAY;
            if (eat(']')) return true;
            for (;;) {
                auto e = std::make_shared<CfgValue>();
                if (!scalar(e.get())) return false;
                if (!v->elems.empty() && e->kind != v->elems[0]->kind)
                    return fail("array elements must share one scalar type");
                v->elems.push_back(e);
                if (eat(']')) return true;
                if (!eat(',')) return fail("',' or ']' expected in array");
            }
        }
        if (eat('(')) {
            v->kind = CfgValue::LIST;
            if (eat(')')) return true;
            for (;;) {
                auto e = std::make_shared<CfgValue>();
                if (!value(e.get())) return false;
                v->elems.push_back(e);
                if (eat(')')) return true;
                if (!eat(',')) return fail("',' or ')' expected in list");
            }
        }
        if (eat('{')) {
            v->kind = CfgValue::GROUP;
            if (!settings(v, '}')) return false;
            return true;
        }
        return scalar(v);
    }
    // settings until the closing character (0: end of text)
    bool settings(CfgValue *g, char close) {
        for (;;) {
            skip();
            if (close ? eat(close) : p >= t.size()) return true;
            if (p >= t.size()) return fail("unexpected end of file");
            std::string n;
            if (!name(&n)) return false;
            if (!eat('=') && !eat(':')) return fail("'=' or ':' expected after '" + n + "'");
            auto v = std::make_shared<CfgValue>();
            if (!value(v.get())) return false;
            if (g->member(n)) return fail("duplicate setting '" + n + "'");
            g->members.emplace_back(n, v);
            if (!eat(';')) eat(',');
        }
    }
};

}  // namespace

bool Config::parse(const std::string &text) {
    Lexer lx(text);
    auto root = std::make_shared<CfgValue>();
    root->kind = CfgValue::GROUP;
    if (!lx.settings(root.get(), 0)) {
        err_ = lx.err;
        root_.reset();
        return false;
    }
    root_ = root;
    err_.clear();
    return true;
}

bool Config::read_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) {
        err_ = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

const CfgValue *Config::lookup(const std::string &path) const {
    const CfgValue *v = root_.get();
    size_t b = 0;
    while (v && b <= path.size()) {
        size_t e = path.find('.', b);
        if (e == std::string::npos) e = path.size();
        const std::string part = path.substr(b, e - b);
        if (part.size() >= 3 && part.front() == '[' && part.back() == ']') {
            if (v->kind != CfgValue::ARRAY && v->kind != CfgValue::LIST) return nullptr;
            char *end = nullptr;
            const long idx = std::strtol(part.c_str() + 1, &end, 10);
            if (*end != ']' || idx < 0 || (size_t)idx >= v->elems.size()) return nullptr;
            v = v->elems[idx].get();
        } else {
            v = v->member(part);
        }
        b = e + 1;
        if (e == path.size()) break;
    }
    return v;
}

bool Config::lookup_bool(const std::string &path, int *out) const {
    const CfgValue *v = lookup(path);
    if (!v || v->kind != CfgValue::BOOL) return false;
    *out = v->b ? 1 : 0;
    return true;
}

bool Config::lookup_int(const std::string &path, int *out) const {
    const CfgValue *v = lookup(path);
    if (!v || v->kind != CfgValue::INT) return false;
    *out = (int)v->i;
    return true;
}

bool Config::lookup_float(const std::string &path, double *out) const {
    const CfgValue *v = lookup(path);
    if (!v || v->kind != CfgValue::FLOAT) return false;
    *out = v->f;
    return true;
}

bool Config::lookup_string(const std::string &path, std::string *out) const {
    const CfgValue *v = lookup(path);
    if (!v || v->kind != CfgValue::STRING) return false;
    *out = v->s;
    return true;
}

bool Config::lookup_strings(const std::string &path, std::vector<std::string> *out) const {
    const CfgValue *v = lookup(path);
    if (!v || (v->kind != CfgValue::ARRAY && v->kind != CfgValue::LIST)) return false;
    out->clear();
    for (const auto &e : v->elems)
        if (e->kind == CfgValue::STRING) out->push_back(e->s);
    return true;
}

}  // namespace spk
