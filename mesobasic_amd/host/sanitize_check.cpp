// sanitize_check.cpp -- exercises the host code of the driver that runs
// without a GPU (the input.dat parser, the VTK writers, the model factory's
// refusal paths) for the AddressSanitizer / UndefinedBehaviorSanitizer build
// (make -C mesobasic_amd/host sanitize).  TEST INFRASTRUCTURE ONLY; exit
// status 0 when every check holds.
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "InputParams.hpp"
#include "LBBaseClass.hpp"
#include "VtkWriter.hpp"

using namespace mesobasic;

namespace {

int failures = 0;

void check(bool ok, const std::string& what)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        failures++;
    }
}

std::string slurp(const std::string& path)
{
    std::ifstream in(path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

void parser_checks()
{
    const char* text =
        "# comment line\n"
        "[Domain]\n nx = 64   ny=32\n nz =16 # trailing\n"
        "[Time]\n nstep = 2000\n dt = 1.0e-0\n"
        "[LBApp]\n app = true\n type = scmp3D\n G = -5.5\n"
        "[./sub]\n key = 'quoted value'\n"
        "[../other]\n x = 3.5\n"
        "[]\n top = 7\n"
        "[Empty]\n"
        "dangling =\n"
        "= lone\n"
        "[\n"
        "]\n"
        "[Output]\n numOutputs = 4 iskip=2\n";
    InputParams p = InputParams::from_string(text);
    check(p("Domain/nx", 0) == 64 && p("Domain/ny", 0) == 32 && p("Domain/nz", 0) == 16, "ints");
    check(p("Time/nstep", 0) == 2000, "nstep");
    check(p("LBApp/G", 0.0) == -5.5, "double");
    check(std::string(p("LBApp/type", "x")) == "scmp3D", "string");
    check(std::string(p("LBApp/sub/key", "x")) == "quoted value", "quoted");
    check(p("other/x", 0.0) == 3.5 || p("LBApp/other/x", 0.0) == 3.5, "parent section");
    check(p("top", 0) == 7, "root section");
    check(p("Missing/key", 42) == 42 && p("Missing/key", 1.5) == 1.5, "defaults");
    check(!p.has("Missing/key"), "has");
    check(p("Output/iskip", 1) == 2, "two keys on a line");
    // numbers GetPot converts through sscanf %lf, ints through double
    InputParams q = InputParams::from_string("a = 12abc\nb = 3.99\nc = -0\nd = 1e400\ne = nan\n");
    check(q("a", 0) == 12, "int prefix");
    check(q("b", 0) == 3, "int via double");
    check(q("c", 1.0) == 0.0, "negative zero");
    (void)q("d", 0.0);
    (void)q("e", 0.0);
    // empty and pathological inputs
    InputParams e = InputParams::from_string("");
    check(e.section_names().empty() && e.variables().empty(), "empty input");
    std::string junk;
    for (int i = 0; i < 2000; i++) junk += (i % 7 == 0) ? "[" : (i % 5 == 0 ? "=" : "k ");
    (void)InputParams::from_string(junk);
    (void)InputParams::from_string(std::string(10000, '='));
    (void)InputParams::from_string("'unterminated quote = 3\n[a]\nb = 1\n");
    bool threw = false;
    try {
        InputParams missing("/nonexistent/dir/input.dat");
    } catch (const std::exception&) {
        threw = true;
    }
    check(threw, "unreadable file throws");
}

void vtk_checks()
{
    char tmpl[] = "/tmp/san_vtk_XXXXXX";
    const char* dir = mkdtemp(tmpl);
    check(dir != nullptr, "mkdtemp");
    if (!dir) return;
    const int NX = 5, NY = 4, NZ = 3;
    std::vector<double> rho(NX * NY * NZ);
    for (size_t i = 0; i < rho.size(); i++) rho[i] = 0.5 + 0.001 * i;
    write_vtk_3d(dir, "rhoA", 7, NX, NY, NZ, 1, 1, 1, rho);
    write_vtk_3d(dir, "rhoA", 8, NX, NY, NZ, 2, 3, 2, rho);
    const std::string f3 = slurp(std::string(dir) + "/rhoA_7.vtk");
    const size_t at = f3.find("DIMENSIONS");
    int dx = 0, dy = 0, dz = 0;
    if (at != std::string::npos) {
        std::istringstream hs(f3.substr(at + 10));
        hs >> dx >> dy >> dz;
    }
    check(dx == NX && dy == NY && dz == NZ, "3-D header");
    std::vector<double> r2(NX * NY), us(NX * NY), vs(NX * NY);
    for (size_t i = 0; i < r2.size(); i++) r2[i] = 1.0 + i, us[i] = 0.01 * i, vs[i] = -0.02 * i;
    write_vtk_2d(dir, "rhoA", 3, NX, NY, 1, 1, r2, us, vs);
    write_vtk_2d(dir, "rhoA", 4, NX, NY, 2, 2, r2, us, vs);
    const std::string f2 = slurp(std::string(dir) + "/rhoA_3.vtk");
    check(f2.find("VECTORS") != std::string::npos, "2-D velocity block");
    for (int s : {7, 8, 3, 4}) std::remove((std::string(dir) + "/rhoA_" + std::to_string(s) + ".vtk").c_str());
    rmdir(dir);
}

void factory_checks()
{
    CommonParams cp{};
    cp.NX = cp.NY = cp.NZ = 8;
    for (const char* t : {"capsule3D", "nosuchtype"}) {
        InputParams in = InputParams::from_string(std::string("[LBApp]\ntype = ") + t + "\n");
        LBBaseClass* m = LBBaseClass::LBFactory(cp, in);
        check(m == nullptr, std::string("factory refuses ") + t);
        delete m;
    }
    // the provided types build a model, or (without a GPU) report why not
    cp.nz = 1;
    for (const char* t : {"scmp2D", "mcmp2D", "mcmp2DFilm", "capsule2D"}) {
        InputParams in = InputParams::from_string(std::string("[LBApp]\ntype = ") + t + "\nseed = 3\n");
        try {
            LBBaseClass* m = LBBaseClass::LBFactory(cp, in);
            check(m != nullptr, std::string("factory builds ") + t);
            delete m;
        } catch (const std::exception& e) {
            check(std::string(e.what()).find("lbm_create") != std::string::npos,
                  std::string("factory error names lbm_create for ") + t);
        }
    }
}

}  // namespace

int main()
{
    parser_checks();
    vtk_checks();
    factory_checks();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("host sanitize check: ok\n");
    return 0;
}
