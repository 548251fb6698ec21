// OpenFOAM ASCII file reading: a lexer over the file text (plain or gzip-compressed, as
// OpenFOAM writes with `writeCompression on`) and a dictionary tree of entries.
//
// The grammar covered is what case files use: `// ...` and `/* ... */` comments, words,
// numbers, "strings", punctuation ( ) { } [ ] ; and `<` `>` inside type words
// (List<scalar>), `N ( ... )` lists, `N{value}` uniform lists, `key value ... ;` entries
// and `key { ... }` sub-dictionaries.  `#include` and `$var` substitutions are not
// supported and are reported as errors when met.  Binary-format files are refused.
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace dugks_foam {

struct Token {
    enum Kind { kWord, kNumber, kPunct, kString, kEnd } kind = kEnd;
    double num = 0.0;   // kNumber
    char punct = 0;     // kPunct
    std::string text;   // kWord / kString (and the literal of a number)
    bool is_int() const;
};

class Lexer {
public:
    explicit Lexer(std::string text) : s_(std::move(text)) {}
    Token next();                 // kEnd at the end of the text
    Token peek();
    size_t line() const { return line_; }

private:
    void skip_space_and_comments();
    std::string s_;
    size_t i_ = 0, line_ = 1;
    bool has_peek_ = false;
    Token peek_;
};

struct Dict;

// One dictionary entry: the tokens up to its ';', or a sub-dictionary.
struct Entry {
    std::vector<Token> tokens;
    std::unique_ptr<Dict> dict;
    bool is_dict() const { return dict != nullptr; }
};

struct Dict {
    std::vector<std::pair<std::string, Entry>> entries;   // file order
    const Entry *find(const std::string &key) const;
    const Dict *sub(const std::string &key) const;         // nullptr when absent or not a dictionary
};

// Reads a file into memory, inflating gzip data; `path` is tried as given and with ".gz".
bool read_file(const std::string &path, std::string &text, std::string &err);

// Parses entries until the end of the text or a closing '}' (which is consumed).
bool parse_dict(Lexer &lx, Dict &d, std::string &err, bool until_brace = false);

// A whole case file: the FoamFile header dictionary and the entries that follow.  When the
// body is not a dictionary (polyMesh lists), `lexer_at_body` leaves the lexer after the
// header for the caller.
struct FoamFile {
    Dict header;
    Dict body;
    std::string cls, object, format;
};
bool read_foam_dict(const std::string &path, FoamFile &f, std::string &err);
bool open_foam_list(const std::string &path, FoamFile &f, std::unique_ptr<Lexer> &lx, std::string &err);

// Token helpers for entries.
bool entry_scalar(const Entry &e, double &v, std::string &err, const std::string &key);
bool entry_word(const Entry &e, std::string &w, std::string &err, const std::string &key);
// `key [dims] value` / `key name [dims] value` / `key value`: the last number of the entry
bool entry_dimensioned(const Entry &e, double &v, std::string &err, const std::string &key);
bool entry_bool(const Entry &e, bool &v, std::string &err, const std::string &key);

// Lists.  A scalar list `N ( a b ... )` or `N{a}`; a vector list `N ( (x y z) ... )`; a
// face list `N ( n(a b c ...) ... )`; read from the lexer at the list's first token.
bool lex_scalar_list(Lexer &lx, std::vector<double> &out, std::string &err);
bool lex_label_list(Lexer &lx, std::vector<int> &out, std::string &err);
bool lex_vector_list(Lexer &lx, std::vector<double> &out, std::string &err);  // [n][3]
bool lex_face_list(Lexer &lx, std::vector<int> &offsets, std::vector<int> &points, std::string &err);

// The same from an entry's tokens starting at token k (advanced past the list).
bool tok_scalar_list(const std::vector<Token> &t, size_t &k, std::vector<double> &out, std::string &err);
bool tok_vector_list(const std::vector<Token> &t, size_t &k, std::vector<double> &out, std::string &err);

}  // namespace dugks_foam
