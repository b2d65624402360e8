// row_parser.cpp -- schema files (the reference's src/attribute.cpp, row.cpp
// and row_parser.cpp:3-28, which read data/<benchmark>.json with a vendored
// JSON library).  This build reads the same layout with a small recursive-
// descent reader of its own: objects, arrays, strings, numbers, true /
// false / null.
#include "row_parser.h"

#include <cstring>
#include <fstream>
#include <sstream>

const std::string &Attribute::getName() const { return name; }
const std::string &Attribute::getType() const { return type; }

void Row::addAttribute(const Attribute &attribute) { attributes.push_back(attribute); }
void Row::set_table_name(const std::string &name) { table_name = name; }
void Row::set_table_size(const int &size) { table_size = size; }
const std::vector<Attribute> &Row::getAttributes() const { return attributes; }
const std::string &Row::get_table_name() const { return table_name; }
const int &Row::get_table_size() const { return table_size; }
void Row::clear() {
    attributes.clear();
    table_name.clear();
    table_size = 0;
}

namespace {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value *get(const std::string &key) const {
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class Reader {
  public:
    Reader(const std::string &text, const std::string &where) : s_(text), where_(where) {}

    Value document() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    [[noreturn]] void fail(const std::string &what) const {
        LOG(FATAL) << "schema " << where_ << ": " << what << " at byte " << i_;
        std::abort();
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool take(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!take(c)) fail(std::string("expected '") + c + "'");
    }
    bool word(const char *w) {
        const size_t n = std::strlen(w);
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    std::string string() {
        expect('"');
        std::string out;
        while (true) {
            if (i_ >= s_.size()) fail("unterminated string");
            char c = s_[i_++];
            if (c == '"') return out;
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (i_ >= s_.size()) fail("bad escape");
            char e = s_[i_++];
            switch (e) {
                case '"': case '\\': case '/': out.push_back(e); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {  // schema names are ASCII; other code points are kept as '?'
                    if (i_ + 4 > s_.size()) fail("bad \\u escape");
                    unsigned v = std::stoul(s_.substr(i_, 4), nullptr, 16);
                    i_ += 4;
                    out.push_back(v < 0x80 ? (char)v : '?');
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            ++i_;
            v.kind = Value::Object;
            if (take('}')) return v;
            do {
                ws();
                std::string k = string();
                expect(':');
                v.obj.emplace_back(std::move(k), value());
            } while (take(','));
            expect('}');
        } else if (c == '[') {
            ++i_;
            v.kind = Value::Array;
            if (take(']')) return v;
            do v.arr.push_back(value());
            while (take(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (word("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (word("false")) {
            v.kind = Value::Bool;
        } else if (word("null")) {
            v.kind = Value::Null;
        } else {
            size_t used = 0;
            try {
                v.num = std::stod(s_.substr(i_, 64), &used);
            } catch (...) {
                fail("expected a value");
            }
            if (used == 0) fail("expected a value");
            i_ += used;
            v.kind = Value::Number;
        }
        return v;
    }

    const std::string &s_;
    std::string where_;
    size_t i_ = 0;
};

const Value &member(const Value &o, const char *key, Value::Kind kind, const std::string &where) {
    const Value *v = o.kind == Value::Object ? o.get(key) : nullptr;
    if (!v || v->kind != kind) LOG(FATAL) << "schema " << where << ": missing or mistyped \"" << key << "\"";
    return *v;
}

}  // namespace

const std::vector<Row> &Row_Parser::parse(const std::string &filename) {
    std::ifstream file(filename);
    if (!file) LOG(FATAL) << "cannot open schema " << filename;
    std::stringstream buf;
    buf << file.rdbuf();
    const std::string text = buf.str();
    const Value doc = Reader(text, filename).document();
    tables.clear();
    uint32_t cur = 0;
    for (const Value &t : member(doc, "tables", Value::Array, filename).arr) {
        Row row;
        row.set_table_name(member(t, "name", Value::String, filename).str);
        row.set_table_size((int)member(t, "size", Value::Number, filename).num);
        for (const Value &a : member(t, "attributes", Value::Array, filename).arr) {
            const std::string &type = member(a, "type", Value::String, filename).str;
            if (type != "int" && type != "double" && type != "string")
                LOG(FATAL) << "schema " << filename << ": table " << row.get_table_name() << ": type " << type;
            const Value *dist = a.get("data_distribution");
            row.addAttribute(Attribute(member(a, "name", Value::String, filename).str, type,
                                       dist && dist->kind == Value::String ? dist->str : std::string()));
            LOG(INFO) << "NO." << cur << "," << row.get_table_name() << "," << row.getAttributes().back().getName()
                    << "," << type;
        }
        ++cur;
        tables.push_back(row);
    }
    return tables;
}
