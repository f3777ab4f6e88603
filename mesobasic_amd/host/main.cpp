// main.cpp -- the meso_amd driver: MesoExecute's process flow
// (src/base/MesoExecute.cpp:12-152) for the lattice-Boltzmann app on one GPU.
//
//   meso_amd [input.dat]            run the simulation in the current directory
//   meso_amd --print-params FILE    print the parameters the run would use
//   meso_amd --serve [options] DIR...
//                                   run every DIR/input.dat (scmp3D, scmp2D)
//                                   together on the batch engine, each
//                                   writing DIR/vtkoutput at its own cadence;
//                                   options: --precision fp32|fp16|bf16|fp16n|
//                                   bf16n (default fp32), --steps-per-tick K,
//                                   --hbm-gib G, --work-limit NODES,
//                                   --max-wait TICKS, --rho-shift X, --quiet
//
// Output: vtkoutput/rhoA_<step>.vtk at step 0 and every nstep/numOutputs
// steps, as the reference writes them; a timing summary on stderr.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mesobasic_lbm.h"
#include "BatchServer.hpp"
#include "InputParams.hpp"
#include "LBApp.hpp"

using namespace mesobasic;

namespace {

void prepare_output_dir()
{
    // MesoExecute.cpp:26-29: mkdir -p vtkoutput; rm -rf vtkoutput/*.vtk
    mkdir("vtkoutput", 0755);
    if (DIR* d = opendir("vtkoutput")) {
        while (dirent* e = readdir(d)) {
            const std::string n = e->d_name;
            if (n.size() > 4 && n.compare(n.size() - 4, 4, ".vtk") == 0)
                std::remove(("vtkoutput/" + n).c_str());
        }
        closedir(d);
    }
}

int print_params(const char* path)
{
    InputParams in(path);
    for (const auto& s : in.section_names()) std::printf("section %s\n", s.c_str());
    for (const auto& kv : in.variables()) std::printf("%s = %s\n", kv.first.c_str(), kv.second.c_str());
    std::printf("nstep %d numOutputs %d\n", in("Time/nstep", 0), in("Output/numOutputs", 1));
    return 0;
}

// The parse of `path` as GetPot reports it (oracle/getpot_check prints the
// same lines from the reference's own GetPot, and a test compares them):
// sections, every variable with its string value, and the typed lookups
// i:KEY:DEFAULT, d:KEY:DEFAULT, s:KEY:DEFAULT.
int query(int argc, char** argv)
{
    InputParams in(argv[2]);
    for (const auto& s : in.section_names()) std::printf("section %s\n", s.c_str());
    for (const auto& kv : in.variables()) std::printf("var %s = [%s]\n", kv.first.c_str(), kv.second.c_str());
    for (int i = 3; i < argc; i++) {
        const std::string q = argv[i];
        const size_t a = q.find(':'), b = q.find(':', a + 1);
        if (a != 1 || b == std::string::npos) throw std::runtime_error("query must be t:key:default");
        const std::string key = q.substr(2, b - 2), def = q.substr(b + 1);
        if (q[0] == 'i') std::printf("i %s %d\n", key.c_str(), in(key.c_str(), std::atoi(def.c_str())));
        else if (q[0] == 'd') std::printf("d %s %.17g\n", key.c_str(), in(key.c_str(), std::strtod(def.c_str(), nullptr)));
        else std::printf("s %s [%s]\n", key.c_str(), in(key.c_str(), def.c_str()).c_str());
    }
    return 0;
}

int serve(int argc, char** argv)
{
    ServeOptions opt;
    std::vector<std::string> dirs;
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        auto value = [&](const char* name) -> std::string {
            if (i + 1 >= argc) throw std::runtime_error(std::string(name) + " needs a value");
            return argv[++i];
        };
        if (a == "--precision") {
            const std::string v = value("--precision");
            if (v == "fp32") opt.precision = LBM_FP32;
            else if (v == "fp16") opt.precision = LBM_FP16S;
            else if (v == "bf16") opt.precision = LBM_BF16S;
            else if (v == "fp16n") opt.precision = LBM_FP16N;
            else if (v == "bf16n") opt.precision = LBM_BF16N;
            else throw std::runtime_error("--precision must be fp32, fp16, bf16, fp16n or bf16n");
        } else if (a == "--steps-per-tick") {
            opt.steps_per_tick = std::stoi(value("--steps-per-tick"));
        } else if (a == "--hbm-gib") {
            opt.hbm_budget = static_cast<int64_t>(std::stod(value("--hbm-gib")) * (1ll << 30));
        } else if (a == "--work-limit") {
            opt.work_limit = std::stoll(value("--work-limit"));
        } else if (a == "--max-wait") {
            opt.max_wait = std::stoi(value("--max-wait"));
        } else if (a == "--rho-shift") {
            opt.rho_shift = std::stod(value("--rho-shift"));
        } else if (a == "--quiet") {
            opt.quiet = true;
        } else if (a.size() > 1 && a[0] == '-' && a[1] == '-') {
            throw std::runtime_error("unknown option " + a);
        } else {
            dirs.push_back(a);
        }
    }
    if (dirs.empty()) throw std::runtime_error("--serve needs at least one run directory");
    return serve_runs(dirs, opt) == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        if (argc == 3 && std::strcmp(argv[1], "--print-params") == 0) return print_params(argv[2]);
        if (argc >= 3 && std::strcmp(argv[1], "--query") == 0) return query(argc, argv);
        if (argc >= 2 && std::strcmp(argv[1], "--serve") == 0) return serve(argc, argv);
        const char* path = argc > 1 ? argv[1] : "input.dat";
        InputParams in(path);
        prepare_output_dir();

        // MesoExecute::createMesoObjects (MesoExecute.cpp:51-99)
        std::vector<std::unique_ptr<MesoBase>> apps;
        for (const std::string& sec : in.section_names()) {
            const std::string flag = in((sec + "app").c_str(), "false");
            if (flag != "true") continue;
            if (sec == "LBApp/") apps.emplace_back(new LBApp(in));
            else
                throw std::runtime_error("section " + sec +
                                         " asks for an app this build does not provide (LBApp only)");
        }
        if (apps.empty()) throw std::runtime_error("no section with app = true");
        for (auto& a : apps) {
            a->initSystem();
            a->writeOutput(0);
        }

        // MesoExecute::executeMesoSimulation (MesoExecute.cpp:107-152)
        const int nstep = in("Time/nstep", 0);
        const int numOutputs = in("Output/numOutputs", 1);
        int outInterval = (numOutputs != 0) ? nstep / numOutputs : nstep + 1;
        if (outInterval == 0) {
            // the reference takes step % 0 here (undefined); write every step
            std::fprintf(stderr, "meso_amd: numOutputs > nstep, writing output every step\n");
            outInterval = 1;
        }
        const auto t0 = std::chrono::steady_clock::now();
        for (int step = 1; step < nstep + 1; step++) {
            for (auto& a : apps) a->stepForward(step);
            if (step % outInterval == 0)
                for (auto& a : apps) a->writeOutput(step);
        }
        for (auto& a : apps) a->finishOutput();
        const double secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        long long nodes = 0;
        for (auto& a : apps) {
            const auto& p = static_cast<LBApp*>(a.get())->params();
            nodes += static_cast<long long>(p.NX) * p.NY * p.NZ;
        }
        std::fprintf(stderr, "meso_amd: %d steps in %.3f s (%.1f MLUPS incl. output)\n", nstep,
                     secs, secs > 0 ? nodes * double(nstep) / secs / 1e6 : 0.0);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "meso_amd: %s\n", e.what());
        return 1;
    }
}
