// OpenFOAM ASCII lexer, dictionary parser and list readers (see foam_dict.h).
#include "foam_dict.h"

#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace dugks_foam {

bool Token::is_int() const { return kind == kNumber && std::floor(num) == num && std::fabs(num) < 2.0e9; }

namespace {
bool is_punct(char c) { return c == '(' || c == ')' || c == '{' || c == '}' || c == '[' || c == ']' || c == ';'; }
}  // namespace

void Lexer::skip_space_and_comments()
{
    const size_t n = s_.size();
    while (i_ < n) {
        const char c = s_[i_];
        if (c == '\n') {
            ++line_;
            ++i_;
        } else if (std::isspace((unsigned char)c)) {
            ++i_;
        } else if (c == '/' && i_ + 1 < n && s_[i_ + 1] == '/') {
            while (i_ < n && s_[i_] != '\n') ++i_;
        } else if (c == '/' && i_ + 1 < n && s_[i_ + 1] == '*') {
            i_ += 2;
            while (i_ + 1 < n && !(s_[i_] == '*' && s_[i_ + 1] == '/')) {
                if (s_[i_] == '\n') ++line_;
                ++i_;
            }
            i_ = std::min(n, i_ + 2);
        } else {
            break;
        }
    }
}

Token Lexer::peek()
{
    if (!has_peek_) {
        peek_ = next();
        has_peek_ = true;
    }
    return peek_;
}

Token Lexer::next()
{
    if (has_peek_) {
        has_peek_ = false;
        return peek_;
    }
    skip_space_and_comments();
    Token t;
    const size_t n = s_.size();
    if (i_ >= n) return t;
    const char c = s_[i_];
    if (is_punct(c)) {
        t.kind = Token::kPunct;
        t.punct = c;
        ++i_;
        return t;
    }
    if (c == '"') {
        size_t j = i_ + 1;
        while (j < n && !(s_[j] == '"' && s_[j - 1] != '\\')) ++j;
        t.kind = Token::kString;
        t.text = s_.substr(i_ + 1, j - i_ - 1);
        i_ = std::min(n, j + 1);
        return t;
    }
    const bool numeric_start = std::isdigit((unsigned char)c) ||
                               ((c == '-' || c == '+' || c == '.') && i_ + 1 < n &&
                                (std::isdigit((unsigned char)s_[i_ + 1]) || s_[i_ + 1] == '.'));
    if (numeric_start) {
        char *end = nullptr;
        const double v = std::strtod(s_.c_str() + i_, &end);
        const size_t len = end ? (size_t)(end - (s_.c_str() + i_)) : 0;
        if (len > 0) {
            t.kind = Token::kNumber;
            t.num = v;
            t.text = s_.substr(i_, len);
            i_ += len;
            return t;
        }
    }
    size_t j = i_;
    while (j < n && !std::isspace((unsigned char)s_[j]) && !is_punct(s_[j]) && s_[j] != '"' &&
           !(s_[j] == '/' && j + 1 < n && (s_[j + 1] == '/' || s_[j + 1] == '*')))
        ++j;
    t.kind = Token::kWord;
    t.text = s_.substr(i_, j - i_);
    i_ = j;
    return t;
}

const Entry *Dict::find(const std::string &key) const
{
    // the last definition wins, as in OpenFOAM
    for (auto it = entries.rbegin(); it != entries.rend(); ++it)
        if (it->first == key) return &it->second;
    return nullptr;
}

const Dict *Dict::sub(const std::string &key) const
{
    const Entry *e = find(key);
    return e && e->is_dict() ? e->dict.get() : nullptr;
}

bool read_file(const std::string &path, std::string &text, std::string &err)
{
    std::string p = path;
    gzFile f = gzopen(p.c_str(), "rb");
    if (!f) {
        p = path + ".gz";
        f = gzopen(p.c_str(), "rb");
    }
    if (!f) {
        err = "cannot open " + path + " (nor " + path + ".gz)";
        return false;
    }
    text.clear();
    char buf[1 << 16];
    for (;;) {
        const int k = gzread(f, buf, sizeof buf);
        if (k < 0) {
            err = "read error in " + p;
            gzclose(f);
            return false;
        }
        if (k == 0) break;
        text.append(buf, (size_t)k);
    }
    gzclose(f);
    return true;
}

bool parse_dict(Lexer &lx, Dict &d, std::string &err, bool until_brace)
{
    for (;;) {
        Token t = lx.next();
        if (t.kind == Token::kEnd) {
            if (until_brace) {
                err = "unexpected end of file inside a dictionary";
                return false;
            }
            return true;
        }
        if (t.kind == Token::kPunct && t.punct == '}') {
            if (until_brace) return true;
            err = "unbalanced '}' at line " + std::to_string(lx.line());
            return false;
        }
        if (t.kind == Token::kPunct && t.punct == ';') continue;  // stray separators
        if (t.kind != Token::kWord && t.kind != Token::kString) {
            err = "expected a keyword at line " + std::to_string(lx.line());
            return false;
        }
        if (!t.text.empty() && (t.text[0] == '#' || t.text[0] == '$')) {
            err = "directive or substitution '" + t.text + "' at line " + std::to_string(lx.line()) +
                  " is not supported";
            return false;
        }
        Entry e;
        Token p = lx.peek();
        if (p.kind == Token::kPunct && p.punct == '{') {
            lx.next();
            e.dict = std::make_unique<Dict>();
            if (!parse_dict(lx, *e.dict, err, true)) return false;
        } else {
            int depth = 0;
            for (;;) {
                Token v = lx.next();
                if (v.kind == Token::kEnd) {
                    err = "entry '" + t.text + "' is not closed by ';'";
                    return false;
                }
                if (v.kind == Token::kPunct) {
                    if (v.punct == '(' || v.punct == '[' || v.punct == '{') ++depth;
                    if (v.punct == ')' || v.punct == ']' || v.punct == '}') --depth;
                    if (depth < 0) {
                        err = "unbalanced bracket in entry '" + t.text + "'";
                        return false;
                    }
                    if (v.punct == ';' && depth == 0) break;
                }
                e.tokens.push_back(std::move(v));
            }
        }
        d.entries.emplace_back(t.text, std::move(e));
    }
}

namespace {
bool read_header(Lexer &lx, FoamFile &f, const std::string &path, std::string &err)
{
    Token t = lx.peek();
    if (t.kind == Token::kWord && t.text == "FoamFile") {
        lx.next();
        Token b = lx.next();
        if (!(b.kind == Token::kPunct && b.punct == '{')) {
            err = path + ": FoamFile header is not a dictionary";
            return false;
        }
        if (!parse_dict(lx, f.header, err, true)) {
            err = path + ": " + err;
            return false;
        }
        std::string e2;
        if (const Entry *c = f.header.find("class")) entry_word(*c, f.cls, e2, "class");
        if (const Entry *o = f.header.find("object")) entry_word(*o, f.object, e2, "object");
        if (const Entry *fm = f.header.find("format")) entry_word(*fm, f.format, e2, "format");
        if (f.format == "binary") {
            err = path + ": binary-format OpenFOAM files are not supported (write the case with writeFormat ascii)";
            return false;
        }
    }
    return true;
}
}  // namespace

bool read_foam_dict(const std::string &path, FoamFile &f, std::string &err)
{
    std::string text;
    if (!read_file(path, text, err)) return false;
    Lexer lx(std::move(text));
    if (!read_header(lx, f, path, err)) return false;
    if (!parse_dict(lx, f.body, err)) {
        err = path + ": " + err;
        return false;
    }
    return true;
}

bool open_foam_list(const std::string &path, FoamFile &f, std::unique_ptr<Lexer> &lx, std::string &err)
{
    std::string text;
    if (!read_file(path, text, err)) return false;
    lx = std::make_unique<Lexer>(std::move(text));
    return read_header(*lx, f, path, err);
}

bool entry_scalar(const Entry &e, double &v, std::string &err, const std::string &key)
{
    if (e.is_dict() || e.tokens.size() != 1 || e.tokens[0].kind != Token::kNumber) {
        err = "entry '" + key + "' is not a single number";
        return false;
    }
    v = e.tokens[0].num;
    return true;
}

bool entry_word(const Entry &e, std::string &w, std::string &err, const std::string &key)
{
    if (e.is_dict() || e.tokens.empty() ||
        (e.tokens[0].kind != Token::kWord && e.tokens[0].kind != Token::kString && e.tokens[0].kind != Token::kNumber)) {
        err = "entry '" + key + "' is not a word";
        return false;
    }
    w = e.tokens[0].text;
    return true;
}

bool entry_dimensioned(const Entry &e, double &v, std::string &err, const std::string &key)
{
    if (e.is_dict()) {
        err = "entry '" + key + "' is a dictionary, not a value";
        return false;
    }
    int depth = 0;
    bool found = false;
    for (const Token &t : e.tokens) {
        if (t.kind == Token::kPunct && t.punct == '[') ++depth;
        else if (t.kind == Token::kPunct && t.punct == ']') --depth;
        else if (depth == 0 && t.kind == Token::kNumber) {
            v = t.num;
            found = true;
        }
    }
    if (!found) err = "entry '" + key + "' has no value";
    return found;
}

bool entry_bool(const Entry &e, bool &v, std::string &err, const std::string &key)
{
    std::string w;
    if (!entry_word(e, w, err, key)) return false;
    if (w == "yes" || w == "true" || w == "on" || w == "y" || w == "t" || w == "1") v = true;
    else if (w == "no" || w == "false" || w == "off" || w == "n" || w == "f" || w == "0" || w == "none") v = false;
    else {
        err = "entry '" + key + "' is not a switch: " + w;
        return false;
    }
    return true;
}

// ----------------------------------------------------------------------------- lists
namespace {
struct LexSrc {
    Lexer &lx;
    Token next() { return lx.next(); }
    Token peek() { return lx.peek(); }
};
struct VecSrc {
    const std::vector<Token> &t;
    size_t &k;
    Token next() { return k < t.size() ? t[k++] : Token{}; }
    Token peek() { return k < t.size() ? t[k] : Token{}; }
};

bool is_p(const Token &t, char c) { return t.kind == Token::kPunct && t.punct == c; }

// one element: a number, or a parenthesised tuple of `width` numbers
template <class Src>
bool element(Src &s, int width, double *out, std::string &err)
{
    if (width == 1) {
        Token t = s.next();
        if (t.kind != Token::kNumber) {
            err = "expected a number in a list, got '" + t.text + "'";
            return false;
        }
        out[0] = t.num;
        return true;
    }
    if (!is_p(s.next(), '(')) {
        err = "expected '(' of a vector";
        return false;
    }
    for (int k = 0; k < width; ++k) {
        Token t = s.next();
        if (t.kind != Token::kNumber) {
            err = "expected a vector component";
            return false;
        }
        out[k] = t.num;
    }
    if (!is_p(s.next(), ')')) {
        err = "expected ')' of a vector";
        return false;
    }
    return true;
}

// `N ( e ... )`, `N{e}`, or `( e ... )` of elements of `width` numbers
template <class Src>
bool list(Src &s, int width, std::vector<double> &out, std::string &err)
{
    out.clear();
    Token t = s.peek();
    long n = -1;
    if (t.kind == Token::kNumber) {
        s.next();
        if (!t.is_int() || t.num < 0) {
            err = "list size is not a non-negative integer";
            return false;
        }
        n = (long)t.num;
        Token b = s.peek();
        if (is_p(b, '{')) {  // uniform list N{e}
            s.next();
            double e[3];
            if (!element(s, width, e, err)) return false;
            if (!is_p(s.next(), '}')) {
                err = "expected '}' of a uniform list";
                return false;
            }
            out.reserve((size_t)n * width);
            for (long i = 0; i < n; ++i) out.insert(out.end(), e, e + width);
            return true;
        }
    }
    if (!is_p(s.next(), '(')) {
        err = "expected '(' of a list";
        return false;
    }
    if (n >= 0) out.reserve((size_t)n * width);
    for (;;) {
        if (is_p(s.peek(), ')')) {
            s.next();
            break;
        }
        if (s.peek().kind == Token::kEnd) {
            err = "list is not closed";
            return false;
        }
        double e[3];
        if (!element(s, width, e, err)) return false;
        out.insert(out.end(), e, e + width);
    }
    if (n >= 0 && (long)(out.size() / width) != n) {
        err = "list announces " + std::to_string(n) + " entries and holds " + std::to_string(out.size() / width);
        return false;
    }
    return true;
}
}  // namespace

bool lex_scalar_list(Lexer &lx, std::vector<double> &out, std::string &err)
{
    LexSrc s{lx};
    return list(s, 1, out, err);
}

bool lex_label_list(Lexer &lx, std::vector<int> &out, std::string &err)
{
    std::vector<double> v;
    LexSrc s{lx};
    if (!list(s, 1, v, err)) return false;
    out.resize(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        if (std::floor(v[i]) != v[i]) {
            err = "label list holds a non-integer";
            return false;
        }
        out[i] = (int)v[i];
    }
    return true;
}

bool lex_vector_list(Lexer &lx, std::vector<double> &out, std::string &err)
{
    LexSrc s{lx};
    return list(s, 3, out, err);
}

bool lex_face_list(Lexer &lx, std::vector<int> &offsets, std::vector<int> &points, std::string &err)
{
    Token t = lx.next();
    if (!t.is_int()) {
        err = "face list size missing";
        return false;
    }
    const long n = (long)t.num;
    if (!is_p(lx.next(), '(')) {
        err = "expected '(' of the face list";
        return false;
    }
    offsets.assign(1, 0);
    offsets.reserve(n + 1);
    for (long f = 0; f < n; ++f) {
        Token k = lx.next();
        if (!k.is_int() || k.num < 3) {
            err = "face " + std::to_string(f) + " has no valid point count";
            return false;
        }
        if (!is_p(lx.next(), '(')) {
            err = "expected '(' of face " + std::to_string(f);
            return false;
        }
        for (int j = 0; j < (int)k.num; ++j) {
            Token p = lx.next();
            if (!p.is_int() || p.num < 0) {
                err = "face " + std::to_string(f) + " holds an invalid point label";
                return false;
            }
            points.push_back((int)p.num);
        }
        if (!is_p(lx.next(), ')')) {
            err = "face " + std::to_string(f) + " is not closed";
            return false;
        }
        offsets.push_back((int)points.size());
    }
    if (!is_p(lx.next(), ')')) {
        err = "face list is not closed";
        return false;
    }
    return true;
}

bool tok_scalar_list(const std::vector<Token> &t, size_t &k, std::vector<double> &out, std::string &err)
{
    VecSrc s{t, k};
    return list(s, 1, out, err);
}

bool tok_vector_list(const std::vector<Token> &t, size_t &k, std::vector<double> &out, std::string &err)
{
    VecSrc s{t, k};
    return list(s, 3, out, err);
}

}  // namespace dugks_foam
