// jsonc.h -- a small reader for JSON with comments (the scenario files carry
// `//` and `/* */` remarks). Host-only; no exceptions leave it.
//
// The document becomes a tree of Node values. Object members keep their file
// order; a lookup returns the first member of that name. Typed reads take a
// fallback that is used when the member is missing or of another type, which
// is how the scenario loader treats optional fields.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace jsonc {

struct Node {
    enum Kind : uint8_t { NUL, BOOL, NUM, STR, ARR, OBJ };
    Kind kind = NUL;
    bool b = false;
    double d = 0.0;
    std::string s;
    std::vector<Node> items;                               // ARR
    std::vector<std::pair<std::string, Node>> members;    // OBJ

    const Node* at(std::string_view key) const {
        if (kind != OBJ) return nullptr;
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
    double as_num(double fallback) const { return kind == NUM ? d : fallback; }
    bool as_bool(bool fallback) const { return kind == BOOL ? b : fallback; }
    std::string as_str(std::string_view fallback) const { return kind == STR ? s : std::string(fallback); }

    double num(std::string_view key, double fallback) const {
        const Node* n = at(key);
        return n ? n->as_num(fallback) : fallback;
    }
    bool flag(std::string_view key, bool fallback) const {
        const Node* n = at(key);
        return n ? n->as_bool(fallback) : fallback;
    }
    std::string str(std::string_view key, std::string_view fallback) const {
        const Node* n = at(key);
        return n ? n->as_str(fallback) : std::string(fallback);
    }
    // Elements of an array member (empty when absent or not an array).
    const std::vector<Node>& list(std::string_view key) const {
        static const std::vector<Node> none;
        const Node* n = at(key);
        return n && n->kind == ARR ? n->items : none;
    }
};

class Reader {
public:
    explicit Reader(std::string_view text) : t_(text) {}

    // Parses the first document of the text; false (with `error`) when it is
    // malformed.
    bool read(Node& out, std::string& error) {
        if (!value(out, 0)) {
            error = err_.empty() ? "malformed JSON" : err_;
            return false;
        }
        return true;   // anything after the document is ignored
    }

private:
    static constexpr int MAX_DEPTH = 64;
    std::string_view t_;
    size_t p_ = 0;
    std::string err_;

    bool bad(const char* what) {
        if (err_.empty()) err_ = std::string(what) + " at byte " + std::to_string(p_);
        return false;
    }
    // Whitespace and comments.
    void skip() {
        for (;;) {
            while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) ++p_;
            if (p_ + 1 < t_.size() && t_[p_] == '/' && t_[p_ + 1] == '/') {
                while (p_ < t_.size() && t_[p_] != '\n') ++p_;
                continue;
            }
            if (p_ + 1 < t_.size() && t_[p_] == '/' && t_[p_ + 1] == '*') {
                const size_t end = t_.find("*/", p_ + 2);
                p_ = end == std::string_view::npos ? t_.size() : end + 2;
                continue;
            }
            return;
        }
    }
    bool literal(std::string_view word) {
        if (t_.substr(p_, word.size()) != word) return bad("unknown literal");
        p_ += word.size();
        return true;
    }
    static void utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s += static_cast<char>(cp);
        } else if (cp < 0x800) {
            s += static_cast<char>(0xC0 | (cp >> 6));
            s += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            s += static_cast<char>(0xE0 | (cp >> 12));
            s += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            s += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    bool string(std::string& s) {
        ++p_;   // opening quote
        while (p_ < t_.size()) {
            const char c = t_[p_++];
            if (c == '"') return true;
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p_ >= t_.size()) break;
            const char e = t_[p_++];
            switch (e) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    if (p_ + 4 > t_.size()) return bad("short \\u escape");
                    uint32_t cp = 0;
                    for (int k = 0; k < 4; ++k) {
                        const char h = t_[p_++];
                        cp <<= 4;
                        if (h >= '0' && h <= '9') cp |= static_cast<uint32_t>(h - '0');
                        else if (h >= 'a' && h <= 'f') cp |= static_cast<uint32_t>(h - 'a' + 10);
                        else if (h >= 'A' && h <= 'F') cp |= static_cast<uint32_t>(h - 'A' + 10);
                        else return bad("bad \\u escape");
                    }
                    utf8(s, cp);
                    break;
                }
                default: return bad("bad escape");
            }
        }
        return bad("unterminated string");
    }
    bool number(double& d) {
        const size_t start = p_;
        if (p_ < t_.size() && (t_[p_] == '-' || t_[p_] == '+')) ++p_;
        while (p_ < t_.size()) {
            const char c = t_[p_];
            if ((c >= '0' && c <= '9') || c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') ++p_;
            else break;
        }
        const std::string tok(t_.substr(start, p_ - start));
        char* end = nullptr;
        d = std::strtod(tok.c_str(), &end);
        if (tok.empty() || end != tok.c_str() + tok.size()) return bad("malformed number");
        return true;
    }
    bool value(Node& n, int depth) {
        if (depth > MAX_DEPTH) return bad("nesting too deep");
        skip();
        if (p_ >= t_.size()) return bad("unexpected end");
        const char c = t_[p_];
        if (c == '{') {
            n.kind = Node::OBJ;
            ++p_;
            skip();
            if (p_ < t_.size() && t_[p_] == '}') {
                ++p_;
                return true;
            }
            for (;;) {
                skip();
                if (p_ >= t_.size() || t_[p_] != '"') return bad("expected a member name");
                std::string key;
                if (!string(key)) return false;
                skip();
                if (p_ >= t_.size() || t_[p_] != ':') return bad("expected ':'");
                ++p_;
                Node v;
                if (!value(v, depth + 1)) return false;
                n.members.emplace_back(std::move(key), std::move(v));
                skip();
                if (p_ < t_.size() && t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == '}') {
                    ++p_;
                    return true;
                }
                return bad("expected ',' or '}'");
            }
        }
        if (c == '[') {
            n.kind = Node::ARR;
            ++p_;
            skip();
            if (p_ < t_.size() && t_[p_] == ']') {
                ++p_;
                return true;
            }
            for (;;) {
                Node v;
                if (!value(v, depth + 1)) return false;
                n.items.push_back(std::move(v));
                skip();
                if (p_ < t_.size() && t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == ']') {
                    ++p_;
                    return true;
                }
                return bad("expected ',' or ']'");
            }
        }
        if (c == '"') {
            n.kind = Node::STR;
            return string(n.s);
        }
        if (c == 't') {
            n.kind = Node::BOOL;
            n.b = true;
            return literal("true");
        }
        if (c == 'f') {
            n.kind = Node::BOOL;
            n.b = false;
            return literal("false");
        }
        if (c == 'n') {
            n.kind = Node::NUL;
            return literal("null");
        }
        n.kind = Node::NUM;
        return number(n.d);
    }
};

}  // namespace jsonc
