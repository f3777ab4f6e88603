// getpot_check.cpp -- reads an input file with the reference's own GetPot
// (src/utils/GetPot, included from the reference tree by oracle/Makefile.ref;
// nothing is copied) and prints what the drop-in driver's parser
// (mesobasic_amd/host/InputParams.cpp, `meso_amd --query`) prints for the
// same file: the section names, every variable with its string value, and
// the typed lookups given on the command line as i:KEY:DEFAULT (int),
// d:KEY:DEFAULT (double) and s:KEY:DEFAULT (string).  TEST INFRASTRUCTURE
// ONLY: it pins the driver's input.dat semantics against GetPot itself.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "utils/GetPot"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: getpot_check input.dat [t:key:default ...]\n");
        return 2;
    }
    GetPot gp(argv[1]);
    for (const std::string& s : gp.get_section_names()) std::printf("section %s\n", s.c_str());
    for (const std::string& v : gp.get_variable_names()) {
        const std::string val = gp(v.c_str(), "");   // GetPot returns a std::string here
        std::printf("var %s = [%s]\n", v.c_str(), val.c_str());
    }
    for (int i = 2; i < argc; i++) {
        const std::string q = argv[i];
        const size_t a = q.find(':'), b = q.find(':', a + 1);
        if (a != 1 || b == std::string::npos) return 2;
        const std::string key = q.substr(2, b - 2), def = q.substr(b + 1);
        if (q[0] == 'i') std::printf("i %s %d\n", key.c_str(), gp(key.c_str(), std::atoi(def.c_str())));
        else if (q[0] == 'd') std::printf("d %s %.17g\n", key.c_str(), gp(key.c_str(), std::strtod(def.c_str(), nullptr)));
        else {
            const std::string val = gp(key.c_str(), def.c_str());
            std::printf("s %s [%s]\n", key.c_str(), val.c_str());
        }
    }
    return 0;
}
