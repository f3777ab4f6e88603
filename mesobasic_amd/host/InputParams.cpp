#include "InputParams.hpp"

#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace mesobasic {

InputParams::InputParams(const std::string& path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot read input file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    parse(ss.str());
}

InputParams InputParams::from_string(const std::string& text)
{
    InputParams p;
    p.parse(text);
    return p;
}

namespace {

// GetPot's reading of a file (GetPot::__read_in_stream, __skip_whitespace,
// __get_next_token, __get_string, __get_until_closing_bracket, GetPot:586-
// 753), restated over the file's bytes with the stream's end-of-file state:
//  - before a token, characters for which isspace() holds are skipped, and a
//    '#' comment up to the end of its line; the character right after '#' is
//    always taken as part of the comment (so an empty "#" line comments out
//    the line after it, as in the reference);
//  - a token runs to the next ' ', '\t' or '\n' that no backslash precedes
//    (a '\r' stays in the token); text between unescaped single quotes is
//    spliced in without the quotes; "${" ... "}" is kept whole; an
//    unescaped backslash is dropped;
//  - an empty token ends the reading: the rest of the file is not read;
//  - the token triple  a "=" b  becomes "a=b" (a trailing "=": "a="); no
//    other spacing round '=' is joined.
class Reader {
public:
    explicit Reader(const std::string& t) : text_(t) {}

    std::vector<std::string> tokens()
    {
        std::vector<std::string> brute;
        while (!failed_) {
            skip_whitespace();
            const std::string tok = next_token();
            if (tok.empty() || tok[0] == static_cast<char>(EOF)) break;
            brute.push_back(tok);
        }
        std::vector<std::string> out;
        size_t i1 = 0, i2 = 1, i3 = 2;
        while (i1 < brute.size()) {
            if (i2 < brute.size() && brute[i2] == "=") {
                out.push_back(i3 >= brute.size() ? brute[i1] + "=" : brute[i1] + "=" + brute[i3]);
                i1 = i3 + 1;
                i2 = i3 + 2;
                i3 = i3 + 3;
                continue;
            }
            out.push_back(brute[i1]);
            i1 = i2;
            i2 = i3;
            i3++;
        }
        return out;
    }

private:
    // istream::get / unget: after a read past the end the stream stays
    // failed, every get returns EOF and unget does nothing
    int get()
    {
        if (failed_ || pos_ >= text_.size()) {
            failed_ = true;
            return EOF;
        }
        return static_cast<unsigned char>(text_[pos_++]);
    }
    void unget()
    {
        if (!failed_ && pos_ > 0) pos_--;
    }

    void skip_whitespace()
    {
        int tmp = get();
        do {
            while (std::isspace(tmp)) {
                tmp = get();
                if (failed_) return;
            }
            if (tmp != '#') {
                unget();
                return;
            }
            tmp = get();   // the character after '#' belongs to the comment
            if (failed_) {
                unget();
                return;
            }
            while (true) {
                tmp = get();
                if (failed_) {
                    unget();
                    return;
                }
                if (tmp == '\n') {
                    unget();
                    break;
                }
            }
            tmp = get();
        } while (!failed_);
        unget();
    }

    std::string next_token()
    {
        std::string token;
        int tmp = 0, last = 0;
        while (true) {
            last = tmp;
            tmp = get();
            if (tmp == EOF || ((tmp == ' ' || tmp == '\t' || tmp == '\n') && last != '\\')) return token;
            if (tmp == '\'' && last != '\\') {
                token += quoted();
                continue;
            }
            if (tmp == '{' && last == '$') {
                token += '{' + bracketed();
                continue;
            }
            if (tmp == '$' && last == '\\') {
                token += '$';
                tmp = 0;
                continue;
            }
            if (tmp == '\\' && last != '\\') continue;
            token += static_cast<char>(tmp);
        }
    }

    std::string quoted()
    {
        std::string str;
        int tmp = 0, last = 0;
        while (true) {
            last = tmp;
            tmp = get();
            if (tmp == EOF) return str;
            if (tmp == '\'' && last != '\\') return str;
            if (tmp == '\\' && last != '\\') continue;
            str += static_cast<char>(tmp);
        }
    }

    std::string bracketed()
    {
        std::string str;
        int tmp = 0, last = 0, brackets = 1;
        while (true) {
            last = tmp;
            tmp = get();
            if (tmp == EOF) return str;
            if (tmp == '{' && last == '$') {
                brackets++;
            } else if (tmp == '}') {
                if (--brackets == 0) return str + '}';
            }
            str += static_cast<char>(tmp);
        }
    }

    const std::string& text_;
    size_t pos_ = 0;
    bool failed_ = false;
};

// GetPot expands "${...}" expressions in every argument and section name
// (GetPot::__DBE_expand_string); this reader does not evaluate them and
// refuses a file that has one, rather than read it differently.
void refuse_expressions(const std::string& arg)
{
    for (size_t i = 0; i + 2 < arg.size(); i++)
        if (arg.compare(i, 2, "${") == 0)
            throw std::runtime_error("input.dat uses a GetPot ${...} expression (" + arg +
                                     "), which this reader does not evaluate");
}

// GetPot::__process_section_label (GetPot:755-800), restated.
std::string section_label(std::string sname, std::vector<std::string>& stack)
{
    if (sname.size() >= 2 && sname.compare(0, 2, "./") == 0) {
        sname = sname.substr(2);
    } else if (sname.size() >= 3 && sname.compare(0, 3, "../") == 0) {
        do {
            if (!stack.empty()) stack.pop_back();
            sname = sname.substr(3);
        } while (sname.compare(0, 3, "../") == 0);
    } else {
        stack.clear();
    }
    if (!sname.empty()) {
        size_t i = 0;
        while (i < sname.size()) {
            if (sname[i] == '/') {
                stack.push_back(sname.substr(0, i));
                sname = (i + 1 < sname.size()) ? sname.substr(i + 1) : std::string();
                i = 0;
            } else {
                ++i;
            }
        }
        if (!sname.empty()) stack.push_back(sname);
    }
    std::string section;
    for (const auto& s : stack) section += s + "/";
    return section;
}

}  // namespace

// GetPot::__parse_argument_vector (GetPot:498-575), restated: a token
// "[name]" sets the section; any other token is prefixed with the section;
// then whatever holds a '=' (a section token too, unprefixed) assigns the
// name before its first '=' (an empty name gives the section itself) the
// rest; a later assignment of a name replaces its value in place.
void InputParams::parse(const std::string& text)
{
    std::vector<std::string> stack;
    std::string section;
    for (std::string arg : Reader(text).tokens()) {
        refuse_expressions(arg);
        if (arg.size() > 1 && arg.front() == '[' && arg.back() == ']') {
            section = section_label(arg.substr(1, arg.size() - 2), stack);
            bool seen = false;
            for (const auto& s : sections_) seen = seen || (s == section);
            if (!seen && !section.empty()) sections_.push_back(section);
        } else {
            arg = section + arg;
        }
        const size_t eq = arg.find('=');
        if (eq == std::string::npos) continue;   // words without '=' assign nothing
        const std::string key = arg.substr(0, eq);
        const std::string val = arg.substr(eq + 1);
        bool replaced = false;
        for (auto& kv : vars_)
            if (kv.first == key) kv.second = val, replaced = true;
        if (!replaced) vars_.emplace_back(key, val);
    }
}

const std::string* InputParams::find(const char* key) const
{
    for (const auto& kv : vars_)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

bool InputParams::has(const char* key) const { return find(key) != nullptr; }

double InputParams::operator()(const char* key, double def) const
{
    const std::string* v = find(key);
    if (!v) return def;
    double tmp;
    if (std::sscanf(v->c_str(), "%lf", &tmp) != 1) return def;
    return tmp;
}

int InputParams::operator()(const char* key, int def) const
{
    return static_cast<int>((*this)(key, static_cast<double>(def)));
}

std::string InputParams::operator()(const char* key, const char* def) const
{
    const std::string* v = find(key);
    return v ? *v : std::string(def);
}

}  // namespace mesobasic
