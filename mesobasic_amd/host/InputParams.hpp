// InputParams.hpp -- reader for the reference's input.dat files.
//
// Same grammar and lookup semantics as the GetPot calls the reference makes
// (src/utils/GetPot, used by MesoExecute.cpp:58-116, LBApp.cpp:18-28,
// scmp3D.cpp:21-31), restated from GetPot's reading and parsing code: the
// file is a stream of tokens split at spaces, tabs and newlines, '#' at the
// start of a token comments to the end of the line, "[Name]" opens a section
// ("[./sub]" relative to the current one, "[../sub]" relative to the parent,
// "[]" back to the root), "key = value" or "key=value" assigns Section/key
// (other spacing round '=' does not: "key =v" assigns the section itself),
// and a later assignment of the same key wins.  Numbers convert with "%lf"
// semantics; integers go through double, as GetPot's int conversion does
// (GetPot:802-818).  A missing or unconvertible key yields the caller's
// default.  GetPot's "${...}" expressions are not evaluated: a file with one
// is refused.  tests/test_getpot_parity.py compares this reader with GetPot
// itself on a corpus of inputs.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace mesobasic {

class InputParams {
public:
    InputParams() = default;
    explicit InputParams(const std::string& path);  // throws std::runtime_error if unreadable
    static InputParams from_string(const std::string& text);

    double operator()(const char* key, double def) const;
    int operator()(const char* key, int def) const;
    std::string operator()(const char* key, const char* def) const;
    bool has(const char* key) const;

    // Section names in order of first appearance, each with a trailing '/'
    // (as GetPot::get_section_names returns them).
    const std::vector<std::string>& section_names() const { return sections_; }
    const std::vector<std::pair<std::string, std::string>>& variables() const { return vars_; }

private:
    void parse(const std::string& text);
    const std::string* find(const char* key) const;
    std::vector<std::string> sections_;
    std::vector<std::pair<std::string, std::string>> vars_;
};

}  // namespace mesobasic
