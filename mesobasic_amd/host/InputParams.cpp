#include "InputParams.hpp"

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

std::vector<std::string> tokenize(const std::string& text)
{
    std::vector<std::string> toks;
    std::string cur;
    bool in_comment = false, in_quote = false;
    for (char ch : text) {
        if (in_comment) {
            if (ch == '\n') in_comment = false;
            continue;
        }
        if (in_quote) {
            if (ch == '\'') in_quote = false;
            else cur += ch;
            continue;
        }
        if (ch == '#') {
            in_comment = true;
            if (!cur.empty()) toks.push_back(cur), cur.clear();
            continue;
        }
        if (ch == '\'') {
            in_quote = true;
            continue;
        }
        if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r') {
            if (!cur.empty()) toks.push_back(cur), cur.clear();
            continue;
        }
        cur += ch;
    }
    if (!cur.empty()) toks.push_back(cur);
    // join "a", "=", "b" / "a=", "b" / "a", "=b" into "a=b"
    std::vector<std::string> out;
    for (size_t i = 0; i < toks.size(); i++) {
        std::string t = toks[i];
        if (t == "=" && !out.empty() && i + 1 < toks.size()) {
            out.back() += "=" + toks[++i];
            continue;
        }
        if (t.size() > 1 && t.back() == '=' && t[0] != '[' && i + 1 < toks.size()) {
            out.push_back(t + toks[++i]);
            continue;
        }
        if (t.size() > 1 && t[0] == '=' && !out.empty()) {
            out.back() += t;
            continue;
        }
        out.push_back(t);
    }
    return out;
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

void InputParams::parse(const std::string& text)
{
    std::vector<std::string> stack;
    std::string section;
    for (const std::string& tok : tokenize(text)) {
        if (tok.size() > 1 && tok.front() == '[' && tok.back() == ']') {
            section = section_label(tok.substr(1, tok.size() - 2), stack);
            bool seen = false;
            for (const auto& s : sections_) seen = seen || (s == section);
            if (!seen && !section.empty()) sections_.push_back(section);
            continue;
        }
        const size_t eq = tok.find('=');
        if (eq == std::string::npos || eq == 0) continue;  // bare words are ignored
        const std::string key = section + tok.substr(0, eq);
        const std::string val = tok.substr(eq + 1);
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
