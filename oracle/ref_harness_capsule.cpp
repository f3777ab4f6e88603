// ref_harness_capsule.cpp -- drives the reference's own immersed-boundary
// capsule model, capsule2D (LBTypes/capsule2D.cpp) with IBcapsules2D
// (LBUtils/IBcapsules2D.cpp), compiled in place from the reference tree by
// oracle/Makefile.ref (no reference source is copied), with an injected
// initial density and capsule, and dumps full-precision binary fields.
// TEST INFRASTRUCTURE ONLY.
//
// capsule2D::initLatticeBoltzmann seeds rand() from time() and places the
// capsule at (30, 30) with radius 8 (capsule2D.cpp:84-120), so the harness
// performs its steps itself on the injected density: rho, zero velocity and
// zero forces on the interior, initCircularCapsule(xo, yo, rad), setFtoFeq.
// Then it calls updateLatticeBoltzmann as LBApp::stepForward does
// (LBApp.cpp:98-114).  One rank only: the IB velocity sum is an Allreduce
// whose order depends on the split.
//
// Usage:
//   ref_harness_capsule nx ny tau nnodes kstiff xo yo rad rho0.bin out_prefix step[,step...]
// Writes out_prefix_<step>.bin: rho | u | v | fx | fy (interior nodes, x
// slowest) | f ([node][9]) | IB x | y | vx | vy | fx | fy (nnodes each) |
// ghost-row force fx, fy (rows j = 0 and j = ny+1, i = 1..nx: 2*nx each),
// all float64.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>
#include <mpi.h>

#define private public
#include "lattice_boltzmann/LBTypes/capsule2D.hpp"
#undef private

namespace {

std::vector<double> read_doubles(const char* path, size_t n)
{
    std::vector<double> v(n);
    FILE* fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(double), n, fp) != n) {
        std::fprintf(stderr, "ref_harness_capsule: cannot read %zu doubles from %s\n", n, path);
        MPI::COMM_WORLD.Abort(2);
    }
    std::fclose(fp);
    return v;
}

}  // namespace

int main(int argc, char** argv)
{
    MPI::Init();
    if (argc != 12 || MPI::COMM_WORLD.Get_size() != 1) {
        std::fprintf(stderr, "usage (one rank): ref_harness_capsule nx ny tau nnodes kstiff xo yo rad rho0.bin "
                             "out steps\n");
        MPI::Finalize();
        return 2;
    }
    const int NX = std::atoi(argv[1]), NY = std::atoi(argv[2]);
    const double tau = std::strtod(argv[3], nullptr);
    const int nnodes = std::atoi(argv[4]);
    const double kstiff = std::strtod(argv[5], nullptr);
    const double xo = std::strtod(argv[6], nullptr), yo = std::strtod(argv[7], nullptr);
    const double rad = std::strtod(argv[8], nullptr);
    const std::string out = argv[10];
    std::vector<int> dumps;
    {
        std::stringstream ss(argv[11]);
        std::string tok;
        while (std::getline(ss, tok, ',')) dumps.push_back(std::atoi(tok.c_str()));
    }
    int last = dumps.empty() ? 0 : *std::max_element(dumps.begin(), dumps.end());
    if (argv[11][0] == '+') {   // "+K": K steps, no dump (timing the reference's step)
        last = std::atoi(argv[11] + 1);
        dumps.clear();
    }

    CommonParams p;
    std::memset(&p, 0, sizeof(p));
    p.NX = NX; p.NY = NY; p.NZ = 1;
    p.nx = NX; p.ny = NY; p.nz = 1;
    p.dx = p.dy = p.dz = 1.0; p.dt = 1.0;
    p.iskip = p.jskip = p.kskip = 1;
    p.np = 1; p.rank = 0; p.nbrL = 0; p.nbrR = 0; p.xOff = 0;
    p.LX = NX; p.LY = NY; p.LZ = 1;

    const std::string inp = out + "_input.dat";
    {
        std::ofstream o(inp.c_str());
        o << std::setprecision(17);
        o << "[LBApp]\n  tauA = " << tau << "\n  nnodes = " << nnodes << "\n  kstiff = " << kstiff
          << "\n  stencil = D2Q9\n[]\n";
    }
    GetPot gp(inp.c_str());
    std::remove(inp.c_str());
    capsule2D m(p, gp);

    const std::vector<double> rho0 = read_doubles(argv[9], static_cast<size_t>(NX) * NY);
    for (int i = 1; i < p.nx + 1; i++)
        for (int j = 1; j < p.ny + 1; j++) {
            const int ndx = i * m.deli + j * m.delj;
            m.fA.setRho(ndx, rho0[static_cast<size_t>(i - 1) * NY + (j - 1)]);
            m.fA.setU(ndx, 0.0);
            m.fA.setV(ndx, 0.0);
            m.fA.setFx(ndx, 0.0);
            m.fA.setFy(ndx, 0.0);
        }
    m.c.initCircularCapsule(xo, yo, rad);
    m.fA.setFtoFeq(m.s);

    for (int step = 0; step <= last; step++) {
        if (step > 0) {
            m.setTimeStep(step);
            m.updateLatticeBoltzmann();
        }
        if (std::find(dumps.begin(), dumps.end(), step) == dumps.end()) continue;
        const std::string path = out + "_" + std::to_string(step) + ".bin";
        FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp) MPI::COMM_WORLD.Abort(3);
        const LBfluid2D& fl = m.fA;
        const std::vector<double>* scal[5] = {&fl.r, &fl.u, &fl.v, &fl.fx, &fl.fy};
        for (const auto* a : scal)
            for (int i = 1; i < p.nx + 1; i++)
                for (int j = 1; j < p.ny + 1; j++) std::fwrite(&(*a)[i * m.deli + j * m.delj], 8, 1, fp);
        for (int i = 1; i < p.nx + 1; i++)
            for (int j = 1; j < p.ny + 1; j++) std::fwrite(&fl.f[(i * m.deli + j * m.delj) * 9], 8, 9, fp);
        const std::vector<double>* ib[6] = {&m.c.x, &m.c.y, &m.c.vx, &m.c.vy, &m.c.fx, &m.c.fy};
        for (const auto* a : ib) std::fwrite(a->data(), 8, nnodes, fp);
        for (int comp = 0; comp < 2; comp++) {
            const std::vector<double>& fa = comp == 0 ? fl.fx : fl.fy;
            for (int row : {0, p.ny + 1})
                for (int i = 1; i < p.nx + 1; i++) std::fwrite(&fa[i * m.deli + row * m.delj], 8, 1, fp);
        }
        std::fclose(fp);
    }
    MPI::Finalize();
    return 0;
}
