// Prints the TOML document named on the command line as tagged JSON (every
// leaf {"type": ..., "value": "..."}), or "ERROR: <message>" with exit code 2
// when the document is rejected.  tests/test_config.py compares the output
// with tomli's reading of the same document.
#include "laminocam_toml.hpp"

#include <cstdio>
#include <string>

using namespace laminocam::toml;

static std::string quote(const std::string &s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        if (c == '"') o += "\\\"";
        else if (c == '\\') o += "\\\\";
        else if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", c);
            o += b;
        } else o += (char)c;
    }
    return o + "\"";
}

static std::string leaf(const char *type, const std::string &v) {
    return std::string("{\"type\": \"") + type + "\", \"value\": " + quote(v) + "}";
}

static std::string dump(const Value &v);

static std::string dump_table(const Table &t) {
    std::string o = "{";
    bool first = true;
    for (const std::string &k : t.keys()) {
        if (!first) o += ", ";
        first = false;
        o += quote(k) + ": " + dump(*t.get(k));
    }
    return o + "}";
}

static std::string dump(const Value &v) {
    switch (v.kind()) {
    case Value::STRING: return leaf("string", v.str());
    case Value::DATETIME: return leaf("datetime", v.str());
    case Value::INTEGER: return leaf("integer", std::to_string(v.integer_value()));
    case Value::BOOLEAN: return leaf("bool", v.bool_value() ? "true" : "false");
    case Value::FLOAT: {
        const double f = v.float_value();
        char b[40];
        if (std::isnan(f)) std::snprintf(b, sizeof b, "nan");
        else if (std::isinf(f)) std::snprintf(b, sizeof b, f > 0 ? "inf" : "-inf");
        else std::snprintf(b, sizeof b, "%.17g", f);
        return leaf("float", b);
    }
    case Value::ARRAY: {
        std::string o = "[";
        const Array &a = *v.as_array();
        for (size_t i = 0; i < a.size(); ++i) o += (i ? ", " : "") + dump(a[i]);
        return o + "]";
    }
    case Value::TABLE: return dump_table(*v.as_table());
    default: return "null";
    }
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.toml>\n", argv[0]);
        return 1;
    }
    try {
        const Table t = parse_file(argv[1]);
        std::printf("%s\n", dump_table(t).c_str());
    } catch (const std::exception &e) {
        std::printf("ERROR: %s\n", e.what());
        return 2;
    }
    return 0;
}
