/*
 * laminocam_toml.hpp - a TOML 1.0 reader for the configuration files of the
 * reconstruction programs, header only, C++17.
 *
 * The reference parses its configuration with toml++ (include/config.h:45-58,
 * toml::parse_file); this header gives the same documents the same meaning:
 * tables, arrays of tables, dotted and quoted keys, inline tables, arrays,
 * basic / literal / multi-line strings with escapes, decimal / hex / octal /
 * binary integers, floats (with inf and nan), booleans, and dates and times
 * (kept as their text).  Duplicate keys and redefined tables are errors.
 *
 *   laminocam::toml::Table t = laminocam::toml::parse_file("config.toml");
 *   const laminocam::toml::Value *v = t.get("recon_params");
 *   std::optional<double> tol = (*v)["tol"].value<double>();
 *
 * Node access mirrors toml++: operator[] on a missing key or a non-table
 * yields an empty node, as_table() / as_array() are null when the node is
 * not one, value<T>() is empty when the node does not hold a T (an integer
 * converts to a floating type; a float converts to an integer type only when
 * it is a whole number in range), value_or(d) falls back to d.
 */
#ifndef LAMINOCAM_TOML_HPP
#define LAMINOCAM_TOML_HPP

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace laminocam::toml {

class Value;
using Array = std::vector<Value>;

class Table {
  public:
    // insertion-ordered map (the order keys appear in the document)
    const Value *get(const std::string &k) const;
    Value *get(const std::string &k);
    Value &insert(const std::string &k, Value v);
    bool contains(const std::string &k) const { return get(k) != nullptr; }
    size_t size() const { return keys_.size(); }
    const std::vector<std::string> &keys() const { return keys_; }
    const Value &operator[](const std::string &k) const;

  private:
    std::vector<std::string> keys_;
    std::map<std::string, std::shared_ptr<Value>> map_;
};

struct parse_error : std::runtime_error {
    int line, column;
    std::string desc;
    parse_error(const std::string &d, int l, int c)
        : std::runtime_error(d + " (line " + std::to_string(l) + ", column " + std::to_string(c) + ")"),
          line(l), column(c), desc(d) {}
    const std::string &description() const { return desc; }
};

class Value {
  public:
    enum Kind { NONE, STRING, INTEGER, FLOAT, BOOLEAN, DATETIME, ARRAY, TABLE };
    Value() = default;
    static Value string(std::string s) { Value v; v.kind_ = STRING; v.s_ = std::move(s); return v; }
    static Value datetime(std::string s) { Value v; v.kind_ = DATETIME; v.s_ = std::move(s); return v; }
    static Value integer(int64_t i) { Value v; v.kind_ = INTEGER; v.i_ = i; return v; }
    static Value floating(double f) { Value v; v.kind_ = FLOAT; v.f_ = f; return v; }
    static Value boolean(bool b) { Value v; v.kind_ = BOOLEAN; v.b_ = b; return v; }
    static Value array() { Value v; v.kind_ = ARRAY; v.a_ = std::make_shared<Array>(); return v; }
    static Value table() { Value v; v.kind_ = TABLE; v.t_ = std::make_shared<Table>(); return v; }

    Kind kind() const { return kind_; }
    explicit operator bool() const { return kind_ != NONE; }
    bool is_table() const { return kind_ == TABLE; }
    bool is_array() const { return kind_ == ARRAY; }
    const Table *as_table() const { return kind_ == TABLE ? t_.get() : nullptr; }
    Table *as_table() { return kind_ == TABLE ? t_.get() : nullptr; }
    const Array *as_array() const { return kind_ == ARRAY ? a_.get() : nullptr; }
    Array *as_array() { return kind_ == ARRAY ? a_.get() : nullptr; }
    const std::string &str() const { return s_; }
    int64_t integer_value() const { return i_; }
    double float_value() const { return f_; }
    bool bool_value() const { return b_; }

    const Value &operator[](const std::string &k) const {
        static const Value none;
        const Table *t = as_table();
        if (!t) return none;
        const Value *v = t->get(k);
        return v ? *v : none;
    }
    const Value &operator[](size_t i) const {
        static const Value none;
        const Array *a = as_array();
        return a && i < a->size() ? (*a)[i] : none;
    }

    template <typename T> std::optional<T> value() const {
        if constexpr (std::is_same_v<T, std::string>) {
            if (kind_ == STRING) return s_;
            return std::nullopt;
        } else if constexpr (std::is_same_v<T, bool>) {
            if (kind_ == BOOLEAN) return b_;
            return std::nullopt;
        } else if constexpr (std::is_floating_point_v<T>) {
            if (kind_ == FLOAT) return (T)f_;
            if (kind_ == INTEGER) return (T)i_;
            return std::nullopt;
        } else if constexpr (std::is_integral_v<T>) {
            long double x;
            if (kind_ == INTEGER) x = (long double)i_;
            else if (kind_ == FLOAT && std::isfinite(f_) && std::floor(f_) == f_) x = f_;
            else return std::nullopt;
            if (x < (long double)std::numeric_limits<T>::lowest() || x > (long double)std::numeric_limits<T>::max())
                return std::nullopt;
            return (T)(kind_ == INTEGER ? (T)i_ : (T)f_);
        } else {
            static_assert(sizeof(T) == 0, "value<T>: unsupported type");
        }
    }
    template <typename T> T value_or(T d) const {
        auto v = value<T>();
        return v ? *v : d;
    }
    std::string value_or(const char *d) const { return value_or(std::string(d)); }

    // parser bookkeeping: a table made by a [header] or an inline table is
    // closed to later [header]s and dotted keys
    // (dotted: made by a dotted key; static_array: an array value, not an
    // array of tables)
    bool defined = false, inline_closed = false, dotted = false, static_array = false;

  private:
    Kind kind_ = NONE;
    std::string s_;
    int64_t i_ = 0;
    double f_ = 0.0;
    bool b_ = false;
    std::shared_ptr<Array> a_;
    std::shared_ptr<Table> t_;
};

inline const Value *Table::get(const std::string &k) const {
    auto it = map_.find(k);
    return it == map_.end() ? nullptr : it->second.get();
}
inline Value *Table::get(const std::string &k) {
    auto it = map_.find(k);
    return it == map_.end() ? nullptr : it->second.get();
}
inline Value &Table::insert(const std::string &k, Value v) {
    keys_.push_back(k);
    auto p = std::make_shared<Value>(std::move(v));
    map_[k] = p;
    return *p;
}
inline const Value &Table::operator[](const std::string &k) const {
    static const Value none;
    const Value *v = get(k);
    return v ? *v : none;
}

namespace detail {

class Parser {
  public:
    explicit Parser(const std::string &src) : s_(src) {}

    Table parse() {
        Value root = Value::table();
        Value *cur = &root;
        for (;;) {
            skip_ws_comments_newlines();
            if (eof()) break;
            if (peek() == '[') {
                cur = header(root);
            } else {
                keyval(*cur->as_table(), false);
            }
            end_of_line();
        }
        return *root.as_table();
    }

  private:
    const std::string &s_;
    size_t p_ = 0;
    int line_ = 1;
    size_t line_start_ = 0;

    [[noreturn]] void fail(const std::string &m) const {
        throw parse_error(m, line_, (int)(p_ - line_start_) + 1);
    }
    bool eof() const { return p_ >= s_.size(); }
    char peek(size_t k = 0) const { return p_ + k < s_.size() ? s_[p_ + k] : '\0'; }
    char get() {
        if (eof()) fail("unexpected end of document");
        const char c = s_[p_++];
        if (c == '\n') {
            ++line_;
            line_start_ = p_;
        }
        return c;
    }
    bool starts(const char *t) const { return s_.compare(p_, std::strlen(t), t) == 0; }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t')) ++p_;
    }
    void skip_comment() {
        if (peek() == '#') {
            while (!eof() && peek() != '\n') {
                const unsigned char c = (unsigned char)peek();
                if ((c < 0x20 && c != '\t' && c != '\r') || c == 0x7f) fail("control character in a comment");
                ++p_;
            }
        }
    }
    bool newline() {
        if (peek() == '\n') {
            get();
            return true;
        }
        if (peek() == '\r' && peek(1) == '\n') {
            ++p_;
            get();
            return true;
        }
        return false;
    }
    void skip_ws_comments_newlines() {
        for (;;) {
            skip_ws();
            skip_comment();
            if (!newline()) return;
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (!eof() && !newline()) fail("expected the end of the line");
    }

    static bool bare_char(char c) {
        return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-';
    }

    std::string simple_key() {
        if (peek() == '"') {
            if (starts("\"\"\"")) fail("multi-line strings are not keys");
            return basic_string();
        }
        if (peek() == '\'') {
            if (starts("'''")) fail("multi-line strings are not keys");
            return literal_string();
        }
        const size_t b = p_;
        while (!eof() && bare_char(peek())) ++p_;
        if (p_ == b) fail("expected a key");
        return s_.substr(b, p_ - b);
    }
    std::vector<std::string> dotted_key() {
        std::vector<std::string> k;
        for (;;) {
            skip_ws();
            k.push_back(simple_key());
            skip_ws();
            if (peek() != '.') return k;
            ++p_;
        }
    }

    // the table at key path `k` below `t`, creating implicit tables
    Value *descend(Value *t, const std::string &key, bool from_dotted) {
        Table *tab = t->as_table();
        Value *v = tab->get(key);
        if (!v) {
            Value nv = Value::table();
            return &tab->insert(key, std::move(nv));
        }
        if (v->is_table()) {
            if (v->inline_closed) fail("inline table '" + key + "' cannot be extended");
            if (from_dotted && v->defined) fail("table '" + key + "' was defined by a header");
            return v;
        }
        if (v->is_array() && !from_dotted) {
            if (v->static_array) fail("'" + key + "' is a static array, not an array of tables");
            Array *a = v->as_array();
            if (a->empty() || !a->back().is_table()) fail("'" + key + "' is not a table");
            return &a->back();
        }
        fail("key '" + key + "' is already a value");
    }

    void keyval(Table &t, bool in_inline) {
        const std::vector<std::string> k = dotted_key();
        skip_ws();
        if (get() != '=') fail("expected '=' after a key");
        skip_ws();
        // the dotted prefix names tables below t, created as needed
        Table *tab = &t;
        for (size_t i = 0; i + 1 < k.size(); ++i) {
            Value *v = tab->get(k[i]);
            if (!v) {
                Value nv = Value::table();
                v = &tab->insert(k[i], std::move(nv));
            } else if (!v->is_table() || v->inline_closed || v->defined) {
                fail("key '" + k[i] + "' cannot be extended by a dotted key");
            }
            v->dotted = true;   // such a table may not be reopened by a header
            tab = v->as_table();
        }
        (void)in_inline;
        if (tab->get(k.back())) fail("duplicate key '" + k.back() + "'");
        Value v = value();
        tab->insert(k.back(), std::move(v));
    }

    Value *header(Value &root) {
        get();   // '['
        const bool aot = peek() == '[';
        if (aot) get();
        const std::vector<std::string> k = dotted_key();
        skip_ws();
        if (get() != ']') fail("expected ']' to close a table header");
        if (aot && get() != ']') fail("expected ']]' to close an array-of-tables header");
        Value *cur = &root;
        for (size_t i = 0; i + 1 < k.size(); ++i) cur = descend(cur, k[i], false);
        Table *tab = cur->as_table();
        Value *v = tab->get(k.back());
        if (aot) {
            if (!v) {
                Value a = Value::array();
                v = &tab->insert(k.back(), std::move(a));
            } else if (!v->is_array() || v->static_array) {
                fail("'" + k.back() + "' is not an array of tables");
            }
            v->as_array()->push_back(Value::table());
            Value &e = v->as_array()->back();
            e.defined = true;
            return &e;
        }
        if (!v) {
            Value nt = Value::table();
            v = &tab->insert(k.back(), std::move(nt));
        } else if (!v->is_table() || v->defined || v->inline_closed || v->dotted) {
            fail("table '" + k.back() + "' is defined more than once");
        }
        v->defined = true;
        return v;
    }

    Value value() {
        const char c = peek();
        if (c == '"') {
            if (starts("\"\"\"")) return Value::string(ml_basic_string());
            return Value::string(basic_string());
        }
        if (c == '\'') {
            if (starts("'''")) return Value::string(ml_literal_string());
            return Value::string(literal_string());
        }
        if (c == '[') return array();
        if (c == '{') return inline_table();
        if (starts("true") && !bare_char(peek(4))) {
            p_ += 4;
            return Value::boolean(true);
        }
        if (starts("false") && !bare_char(peek(5))) {
            p_ += 5;
            return Value::boolean(false);
        }
        return number_or_date();
    }

    Value array() {
        get();   // '['
        Value a = Value::array();
        a.static_array = true;
        for (;;) {
            skip_ws_comments_newlines();
            if (peek() == ']') {
                get();
                return a;
            }
            a.as_array()->push_back(value());
            skip_ws_comments_newlines();
            if (peek() == ',') {
                get();
                continue;
            }
            if (peek() == ']') {
                get();
                return a;
            }
            fail("expected ',' or ']' in an array");
        }
    }

    Value inline_table() {
        get();   // '{'
        Value t = Value::table();
        skip_ws();
        if (peek() == '}') {
            get();
            t.inline_closed = true;
            return t;
        }
        for (;;) {
            skip_ws();
            keyval(*t.as_table(), true);
            skip_ws();
            const char c = get();
            if (c == '}') break;
            if (c != ',') fail("expected ',' or '}' in an inline table");
        }
        close_inline(t);
        return t;
    }
    static void close_inline(Value &t) {
        t.inline_closed = true;
        if (Table *tab = t.as_table())
            for (const std::string &k : tab->keys()) {
                Value *v = tab->get(k);
                if (v->is_table()) close_inline(*v);
            }
    }

    void utf8(std::string &out, uint32_t cp) {
        if (cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) fail("invalid unicode escape");
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
    void escape(std::string &out) {
        const char e = get();
        switch (e) {
        case 'b': out += '\b'; break;
        case 't': out += '\t'; break;
        case 'n': out += '\n'; break;
        case 'f': out += '\f'; break;
        case 'r': out += '\r'; break;
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case 'u':
        case 'U': {
            const int n = e == 'u' ? 4 : 8;
            uint32_t cp = 0;
            for (int i = 0; i < n; ++i) {
                const char h = get();
                int d;
                if (h >= '0' && h <= '9') d = h - '0';
                else if (h >= 'a' && h <= 'f') d = h - 'a' + 10;
                else if (h >= 'A' && h <= 'F') d = h - 'A' + 10;
                else fail("invalid unicode escape");
                cp = cp * 16 + (uint32_t)d;
            }
            utf8(out, cp);
            break;
        }
        default: fail(std::string("invalid escape '\\") + e + "'");
        }
    }
    void check_char(char c, bool multiline) {
        const unsigned char u = (unsigned char)c;
        if ((u < 0x20 && u != '\t' && !(multiline && (u == '\n' || u == '\r'))) || u == 0x7f)
            fail("control character in a string");
    }
    std::string basic_string() {
        get();
        std::string out;
        for (;;) {
            if (eof() || peek() == '\n') fail("unterminated string");
            const char c = get();
            if (c == '"') return out;
            if (c == '\\') {
                escape(out);
                continue;
            }
            check_char(c, false);
            out += c;
        }
    }
    std::string literal_string() {
        get();
        std::string out;
        for (;;) {
            if (eof() || peek() == '\n') fail("unterminated string");
            const char c = get();
            if (c == '\'') return out;
            check_char(c, false);
            out += c;
        }
    }
    std::string ml_basic_string() {
        p_ += 3;
        newline();   // a newline right after the opening quotes is trimmed
        std::string out;
        for (;;) {
            if (eof()) fail("unterminated multi-line string");
            if (starts("\"\"\"")) {
                // up to two quotes may precede the closing delimiter
                size_t q = 3;
                while (q < 5 && peek(q) == '"') ++q;
                out.append(q - 3, '"');
                p_ += q;
                return out;
            }
            const char c = get();
            if (c == '\\') {
                // line-ending backslash: drop whitespace and newlines up to the next text
                size_t k = p_;
                while (k < s_.size() && (s_[k] == ' ' || s_[k] == '\t')) ++k;
                if (k < s_.size() && (s_[k] == '\n' || (s_[k] == '\r' && k + 1 < s_.size() && s_[k + 1] == '\n'))) {
                    p_ = k;
                    for (;;) {
                        skip_ws();
                        if (!newline()) break;
                    }
                    continue;
                }
                escape(out);
                continue;
            }
            if (c == '\r' && peek() == '\n') continue;
            check_char(c, true);
            out += c;
        }
    }
    std::string ml_literal_string() {
        p_ += 3;
        newline();
        std::string out;
        for (;;) {
            if (eof()) fail("unterminated multi-line string");
            if (starts("'''")) {
                size_t q = 3;
                while (q < 5 && peek(q) == '\'') ++q;
                out.append(q - 3, '\'');
                p_ += q;
                return out;
            }
            const char c = get();
            if (c == '\r' && peek() == '\n') continue;
            check_char(c, true);
            out += c;
        }
    }

    static bool is_digit(char c) { return c >= '0' && c <= '9'; }

    // digits with single underscores between digits
    std::string digits(bool (*ok)(char), const char *what) {
        std::string d;
        bool last_us = true;
        while (!eof() && (ok(peek()) || peek() == '_')) {
            if (peek() == '_') {
                if (last_us) fail(std::string("misplaced '_' in ") + what);
                last_us = true;
                ++p_;
                continue;
            }
            d += get();
            last_us = false;
        }
        if (d.empty() || last_us) fail(std::string("malformed ") + what);
        return d;
    }

    Value number_or_date() {
        // dates and times: 1979-05-27, 07:32:00, 1979-05-27T07:32:00Z ...
        if (is_digit(peek()) && is_digit(peek(1)) &&
            ((is_digit(peek(2)) && is_digit(peek(3)) && peek(4) == '-') || peek(2) == ':')) {
            const size_t b = p_;
            while (!eof()) {
                const char c = peek();
                if (is_digit(c) || c == '-' || c == ':' || c == '.' || c == 'T' || c == 't' || c == 'Z' ||
                    c == 'z' || c == '+') {
                    ++p_;
                } else if (c == ' ' && is_digit(peek(1)) && p_ - b == 10) {
                    ++p_;   // date and time separated by a space
                } else {
                    break;
                }
            }
            return Value::datetime(s_.substr(b, p_ - b));
        }
        const size_t b = p_;
        bool neg = false;
        if (peek() == '+' || peek() == '-') {
            neg = get() == '-';
        }
        if (starts("inf") || starts("nan")) {
            const bool inf = peek() == 'i';
            p_ += 3;
            if (bare_char(peek())) fail("malformed number");
            const double v = inf ? std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
            return Value::floating(neg ? -v : v);
        }
        if (peek() == '0' && (peek(1) == 'x' || peek(1) == 'o' || peek(1) == 'b')) {
            if (p_ != b) fail("a sign is not allowed on hexadecimal, octal or binary integers");
            const char base = peek(1);
            p_ += 2;
            std::string d;
            int radix;
            if (base == 'x') {
                radix = 16;
                d = digits([](char c) { return is_digit(c) || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'); },
                           "hexadecimal integer");
            } else if (base == 'o') {
                radix = 8;
                d = digits([](char c) { return c >= '0' && c <= '7'; }, "octal integer");
            } else {
                radix = 2;
                d = digits([](char c) { return c == '0' || c == '1'; }, "binary integer");
            }
            uint64_t v = 0;
            for (char c : d) {
                const int x = is_digit(c) ? c - '0' : (c | 0x20) - 'a' + 10;
                if (v > (uint64_t)(std::numeric_limits<int64_t>::max() - x) / (uint64_t)radix)
                    fail("integer out of range");
                v = v * (uint64_t)radix + (uint64_t)x;
            }
            if (bare_char(peek()) || peek() == '.') fail("malformed number");
            return Value::integer((int64_t)v);
        }
        const std::string ip = digits(is_digit, "number");
        if (ip.size() > 1 && ip[0] == '0') fail("leading zeros are not allowed");
        std::string text = (neg ? "-" : "") + ip;
        bool is_float = false;
        if (peek() == '.') {
            ++p_;
            text += '.' + digits(is_digit, "fraction");
            is_float = true;
        }
        if (peek() == 'e' || peek() == 'E') {
            ++p_;
            text += 'e';
            if (peek() == '+' || peek() == '-') text += get();
            text += digits(is_digit, "exponent");
            is_float = true;
        }
        if (bare_char(peek()) || peek() == '.') fail("malformed number");
        if (is_float) return Value::floating(std::strtod(text.c_str(), nullptr));
        errno = 0;
        const long long v = std::strtoll(text.c_str(), nullptr, 10);
        if (errno == ERANGE) fail("integer out of range");
        return Value::integer((int64_t)v);
    }
};

} // namespace detail

inline Table parse(const std::string &text) { return detail::Parser(text).parse(); }

inline Table parse_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

} // namespace laminocam::toml

#endif
