// sanitize_check.cpp -- exercises the host code of the driver that runs
// without a GPU (the input.dat parser, the VTK writers, the model factory's
// refusal paths) for the AddressSanitizer / UndefinedBehaviorSanitizer build
// (make -C mesobasic_amd/host sanitize).  TEST INFRASTRUCTURE ONLY; exit
// status 0 when every check holds.
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
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
    // GetPot joins only "a = b" / "a=b": "nz =16" assigns the section itself
    check(p("Domain/nx", 0) == 64 && p("Domain/ny", 0) == 32 && p("Domain/nz", 0) == 0 &&
              p("Domain/", 0) == 16,
          "ints");
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

// put_fixed3 against printf("%.3f"): special values, ties (exactly
// representable values halfway between two 3-decimal results), values next
// to the rounding boundaries, every exponent, and random ones.
void format_checks()
{
    auto same = [](double v) {
        char a[400], b[kMaxFixed3 + 1];
        std::snprintf(a, sizeof a, "%.3f", v);
        char* e = put_fixed3(b, v);
        *e = 0;
        return std::string(a) == std::string(b);
    };
    int bad = 0;
    std::vector<double> vals = {0.0, -0.0, 1.0, -1.0, 0.0005, -0.0005, 0.0015, 0.0025, 0.0625, 0.1875,
                                1.0005, 2.0625, 0.9995, 999.9995, 1e-300, 5e-324, 1e308, -1e308,
                                1.7976931348623157e308, 123456789.0125, 0.7, 2.0, 0.15};
    vals.push_back(std::numeric_limits<double>::quiet_NaN());
    vals.push_back(-std::numeric_limits<double>::quiet_NaN());
    vals.push_back(std::numeric_limits<double>::infinity());
    vals.push_back(-std::numeric_limits<double>::infinity());
    for (int k = -20000; k <= 20000; k++) {
        const double t = (k + 0.5) / 1000.0;   // near the boundaries
        vals.push_back(t);
        vals.push_back(std::nextafter(t, 1e9));
        vals.push_back(std::nextafter(t, -1e9));
        vals.push_back(k / 16.0);               // exact binary ties at 3 decimals
    }
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < 200000; i++) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        double v;
        std::memcpy(&v, &x, sizeof v);         // every exponent
        vals.push_back(v);
        vals.push_back(static_cast<double>(x >> 11) / 9007199254740992.0 * 20.0 - 10.0);
    }
    for (double v : vals) bad += !same(v);
    check(bad == 0, "put_fixed3 equals printf %.3f (" + std::to_string(bad) + " differ)");
}

// The writers' files against the same loops printed through an ostream
// (fixed, 3 decimals, one value per line), on a lattice large enough for
// several formatting threads, with and without skips.
void vtk_bytes_checks()
{
    char tmpl[] = "/tmp/san_vtkb_XXXXXX";
    const char* dir = mkdtemp(tmpl);
    check(dir != nullptr, "mkdtemp");
    if (!dir) return;
    const int NX = 70, NY = 64, NZ = 40;
    std::vector<double> rho(static_cast<size_t>(NX) * NY * NZ);
    uint64_t x = 12345;
    for (auto& r : rho) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        r = 0.1 + 2.0 * static_cast<double>(x >> 11) / 9007199254740992.0;
    }
    rho[7] = -0.0;
    rho[99] = std::numeric_limits<double>::quiet_NaN();
    for (int sk : {1, 3}) {
        const int step = 100 + sk;
        write_vtk_3d(dir, "rhoA", step, NX, NY, NZ, sk, sk, sk, rho);
        const std::string got = slurp(std::string(dir) + "/rhoA_" + std::to_string(step) + ".vtk");
        std::ostringstream want;
        for (int k = 1; k < NZ + 1; k += sk)
            for (int j = 1; j < NY + 1; j += sk)
                for (int i = 1; i < NX + 1; i++)
                    if (i % sk == 0)
                        want << std::fixed << std::setprecision(3)
                             << rho[(static_cast<size_t>(i - 1) * NY + (j - 1)) * NZ + (k - 1)] << "\n";
        const size_t body = got.find("LOOKUP_TABLE default\n");
        check(body != std::string::npos && got.substr(body + 21) == want.str(),
              "3-D VTK body equals the ostream loop, skip " + std::to_string(sk));
        std::remove((std::string(dir) + "/rhoA_" + std::to_string(step) + ".vtk").c_str());
    }
    const int MX = 700, MY = 400;
    std::vector<double> r2(static_cast<size_t>(MX) * MY), us(r2.size()), vs(r2.size());
    for (size_t c = 0; c < r2.size(); c++) r2[c] = rho[c % rho.size()], us[c] = 0.01 * (c % 97) - 0.4, vs[c] = -1e-4 * c;
    write_vtk_2d(dir, "rhoA", 5, MX, MY, 2, 1, r2, us, vs);
    const std::string got = slurp(std::string(dir) + "/rhoA_5.vtk");
    std::ostringstream want;
    for (int j = 1; j < MY + 1; j += 1)
        for (int i = 1; i < MX + 1; i++)
            if (i % 2 == 0) want << std::fixed << std::setprecision(3) << r2[static_cast<size_t>(i - 1) * MY + (j - 1)] << "\n";
    want << "   " << "\n" << "VECTORS Velocity float" << "\n";
    for (int j = 1; j < MY + 1; j += 1)
        for (int i = 1; i < MX + 1; i++)
            if (i % 2 == 0) {
                const size_t c = static_cast<size_t>(i - 1) * MY + (j - 1);
                want << std::fixed << std::setprecision(3) << us[c] << " " << vs[c] << " " << 0.0 << "\n";
            }
    const size_t body = got.find("LOOKUP_TABLE default\n");
    check(body != std::string::npos && got.substr(body + 21) == want.str(), "2-D VTK body equals the ostream loop");
    std::remove((std::string(dir) + "/rhoA_5.vtk").c_str());
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
    format_checks();
    vtk_bytes_checks();
    factory_checks();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("host sanitize check: ok\n");
    return 0;
}
