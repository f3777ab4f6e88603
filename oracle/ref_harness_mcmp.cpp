// ref_harness_mcmp.cpp -- drives the reference's own two-component
// Shan-Chen classes, mcmp2D (LBTypes/mcmp2D.cpp) and mcmp3D
// (LBTypes/mcmp3D.cpp), and their film forms mcmp2DFilm / mcmp3DFilm (walls
// on the last axis, LBTypes/mcmp2DFilm.cpp, mcmp3DFilm.cpp), compiled in
// place from the reference tree by
// oracle/Makefile.ref (no reference source is copied), with injected initial
// densities, and dumps full-precision binary fields.  TEST INFRASTRUCTURE
// ONLY.
//
// The models seed rand() from time() (mcmp2D.cpp:65, mcmp3D.cpp:65), so the
// harness performs initLatticeBoltzmann's remaining steps itself on the
// injected densities (ghost rho of both components, forces and common
// velocity, setFtoFeq of both: mcmp2D.cpp:83-87, mcmp3D.cpp:87-91) and then
// calls updateLatticeBoltzmann as LBApp::stepForward does (LBApp.cpp:98-114).
//
// Usage:
//   ref_harness_mcmp q nx ny nz tauA tauB G rhoA0.bin rhoB0.bin out_prefix step[,step...]
//                    [rhoA_sol rhoB_sol]
// With the two wall densities the film classes run instead (their keys
// LBApp/rhoA_sol, rhoB_sol).  A step list "+K" runs K steps and dumps nothing
// (for timing the reference's step).
// Writes out_prefix_<c>_<step>.bin for c = A, B: rho | u | v | w | fx | fy | fz
// | f (float64, interior nodes, x slowest and the last axis fastest; f is
// [node][q]; w and fz are zeros in 2-D).  Run under mpirun -np N to split x
// into slabs (LBApp.cpp:46-60); results do not depend on N.

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
#include "lattice_boltzmann/LBTypes/mcmp2D.hpp"
#include "lattice_boltzmann/LBTypes/mcmp2DFilm.hpp"
#include "lattice_boltzmann/LBTypes/mcmp3D.hpp"
#include "lattice_boltzmann/LBTypes/mcmp3DFilm.hpp"
#undef private

namespace {

// x slab of this rank, as LBApp::LBApp splits the lattice (LBApp.cpp:29-60)
CommonParams slab_params(int NX, int NY, int NZ)
{
    CommonParams p;
    std::memset(&p, 0, sizeof(p));
    p.NX = NX; p.NY = NY; p.NZ = NZ;
    p.dx = p.dy = p.dz = 1.0; p.dt = 1.0;
    p.iskip = p.jskip = p.kskip = 1;
    p.np = MPI::COMM_WORLD.Get_size();
    p.rank = MPI::COMM_WORLD.Get_rank();
    p.nbrL = (p.rank - 1) + ((p.rank - 1) < 0) * p.np;
    p.nbrR = (p.rank + 1) - ((p.rank + 1) > (p.np - 1)) * p.np;
    const int base = NX / p.np, extra = NX % p.np;
    p.nx = (p.rank < extra) ? base + 1 : base;
    int off = 0;
    for (int r = 0; r < p.rank; r++) off += (r < extra) ? base + 1 : base;
    p.ny = NY; p.nz = NZ; p.xOff = off;
    p.LX = NX; p.LY = NY; p.LZ = NZ;
    return p;
}

std::vector<double> read_doubles(const char* path, size_t n)
{
    std::vector<double> v(n);
    FILE* fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(double), n, fp) != n) {
        std::fprintf(stderr, "ref_harness_mcmp: cannot read %zu doubles from %s\n", n, path);
        MPI::COMM_WORLD.Abort(2);
    }
    std::fclose(fp);
    return v;
}

// Local block of one rank: nscal scalar fields of nloc values, then f
// (nloc * q).  Rank 0 gathers the slabs (x slowest, so each field is the
// rank-ordered concatenation) and writes field by field.
void gather_write(const std::string& path, const std::vector<double>& blk, long nloc, int nscal, int q)
{
    const int np = MPI::COMM_WORLD.Get_size(), rank = MPI::COMM_WORLD.Get_rank();
    int cnt = static_cast<int>(blk.size());
    std::vector<int> counts(np), displs(np);
    MPI::COMM_WORLD.Gather(&cnt, 1, MPI::INT, counts.data(), 1, MPI::INT, 0);
    std::vector<double> all;
    if (rank == 0) {
        int tot = 0;
        for (int r = 0; r < np; r++) { displs[r] = tot; tot += counts[r]; }
        all.resize(tot);
    }
    MPI::COMM_WORLD.Gatherv(blk.data(), cnt, MPI::DOUBLE, all.data(), counts.data(), displs.data(),
                            MPI::DOUBLE, 0);
    (void)nloc;
    if (rank != 0) return;
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) MPI::COMM_WORLD.Abort(3);
    for (int field = 0; field <= nscal; field++)
        for (int r = 0; r < np; r++) {
            const long nr = counts[r] / (nscal + q);
            const double* b = all.data() + displs[r];
            if (field < nscal) std::fwrite(b + field * nr, sizeof(double), nr, fp);
            else std::fwrite(b + nscal * nr, sizeof(double), nr * q, fp);
        }
    std::fclose(fp);
}

std::string input_file(const std::string& out, int rank, const char* stencil, double tauA, double tauB,
                       double G, const double* sol)
{
    char name[512];
    std::snprintf(name, sizeof(name), "%s_rank%d_input.dat", out.c_str(), rank);
    std::ofstream o(name);
    o << std::setprecision(17);
    o << "[LBApp]\n  G = " << G << "\n  tauA = " << tauA << "\n  tauB = " << tauB
      << "\n  stencil = " << stencil << "\n";
    if (sol) o << "  rhoA_sol = " << sol[0] << "\n  rhoB_sol = " << sol[1] << "\n";
    o << "[]\n";
    return name;
}

struct Run {
    CommonParams p;
    int NX, NY, NZ;
    std::vector<double> rA, rB;
    std::vector<int> dumps;
    int last;
    std::string out;
};

// initLatticeBoltzmann minus the rand() fill, then updateLatticeBoltzmann
// per step; dumps both components at the listed steps.
template <typename M>
void run3d(M& m, const Run& r)
{
    const CommonParams& p = r.p;
    const int nscal = 7;
    for (int i = 1; i < p.nx + 1; i++)
        for (int j = 1; j < p.ny + 1; j++)
            for (int k = 1; k < p.nz + 1; k++) {
                const int ndx = i * m.deli + j * m.delj + k * m.delk;
                const long g = ((long)(i - 1 + p.xOff) * r.NY + (j - 1)) * r.NZ + (k - 1);
                m.fA.setRho(ndx, r.rA[g]);
                m.fB.setRho(ndx, r.rB[g]);
                m.fA.setU(ndx, 0.0); m.fA.setV(ndx, 0.0); m.fA.setW(ndx, 0.0);
                m.fB.setU(ndx, 0.0); m.fB.setV(ndx, 0.0); m.fB.setW(ndx, 0.0);
            }
    m.fA.ghostNodesRho();
    m.fB.ghostNodesRho();
    m.calculateShanChenForces();
    m.fA.setFtoFeq(m.s);
    m.fB.setFtoFeq(m.s);
    for (int step = 0; step <= r.last; step++) {
        if (step > 0) {
            m.setTimeStep(step);
            m.updateLatticeBoltzmann();
        }
        if (std::find(r.dumps.begin(), r.dumps.end(), step) == r.dumps.end()) continue;
        const long nloc = (long)p.nx * p.ny * p.nz;
        for (int comp = 0; comp < 2; comp++) {
            const LBfluid3D& fl = comp == 0 ? m.fA : m.fB;
            std::vector<double> blk((nscal + 19) * nloc, 0.0);
            long c = 0;
            for (int i = 1; i < p.nx + 1; i++)
                for (int j = 1; j < p.ny + 1; j++)
                    for (int k = 1; k < p.nz + 1; k++, c++) {
                        const long ndx = i * m.deli + j * m.delj + k * m.delk;
                        blk[0 * nloc + c] = fl.r[ndx];
                        blk[1 * nloc + c] = fl.u[ndx];
                        blk[2 * nloc + c] = fl.v[ndx];
                        blk[3 * nloc + c] = fl.w[ndx];
                        blk[4 * nloc + c] = fl.fx[ndx];
                        blk[5 * nloc + c] = fl.fy[ndx];
                        blk[6 * nloc + c] = fl.fz[ndx];
                        for (int n = 0; n < 19; n++) blk[7 * nloc + c * 19 + n] = fl.f[ndx * 19 + n];
                    }
            gather_write(r.out + (comp == 0 ? "_A_" : "_B_") + std::to_string(step) + ".bin", blk, nloc,
                         nscal, 19);
        }
    }
}

template <typename M>
void run2d(M& m, const Run& r)
{
    const CommonParams& p = r.p;
    const int nscal = 7;
    for (int i = 1; i < p.nx + 1; i++)
        for (int j = 1; j < p.ny + 1; j++) {
            const int ndx = i * m.deli + j * m.delj;
            const long g = (long)(i - 1 + p.xOff) * r.NY + (j - 1);
            m.fA.setRho(ndx, r.rA[g]);
            m.fB.setRho(ndx, r.rB[g]);
            m.fA.setU(ndx, 0.0); m.fA.setV(ndx, 0.0);
            m.fB.setU(ndx, 0.0); m.fB.setV(ndx, 0.0);
        }
    m.fA.ghostNodesRho();
    m.fB.ghostNodesRho();
    m.calculateShanChenForces();
    m.fA.setFtoFeq(m.s);
    m.fB.setFtoFeq(m.s);
    for (int step = 0; step <= r.last; step++) {
        if (step > 0) {
            m.setTimeStep(step);
            m.updateLatticeBoltzmann();
        }
        if (std::find(r.dumps.begin(), r.dumps.end(), step) == r.dumps.end()) continue;
        const long nloc = (long)p.nx * p.ny;
        for (int comp = 0; comp < 2; comp++) {
            const LBfluid2D& fl = comp == 0 ? m.fA : m.fB;
            std::vector<double> blk((nscal + 9) * nloc, 0.0);
            long c = 0;
            for (int i = 1; i < p.nx + 1; i++)
                for (int j = 1; j < p.ny + 1; j++, c++) {
                    const long ndx = i * m.deli + j * m.delj;
                    blk[0 * nloc + c] = fl.r[ndx];
                    blk[1 * nloc + c] = fl.u[ndx];
                    blk[2 * nloc + c] = fl.v[ndx];
                    blk[4 * nloc + c] = fl.fx[ndx];
                    blk[5 * nloc + c] = fl.fy[ndx];
                    for (int n = 0; n < 9; n++) blk[7 * nloc + c * 9 + n] = fl.f[ndx * 9 + n];
                }
            gather_write(r.out + (comp == 0 ? "_A_" : "_B_") + std::to_string(step) + ".bin", blk, nloc,
                         nscal, 9);
        }
    }
}

}  // namespace

int main(int argc, char** argv)
{
    MPI::Init();
    if (argc != 12 && argc != 14) {
        if (MPI::COMM_WORLD.Get_rank() == 0)
            std::fprintf(stderr, "usage: ref_harness_mcmp q nx ny nz tauA tauB G rhoA0.bin rhoB0.bin out steps "
                                 "[rhoA_sol rhoB_sol]\n");
        MPI::Finalize();
        return 2;
    }
    Run r;
    const int q = std::atoi(argv[1]);
    r.NX = std::atoi(argv[2]); r.NY = std::atoi(argv[3]); r.NZ = std::atoi(argv[4]);
    const double tauA = std::strtod(argv[5], nullptr), tauB = std::strtod(argv[6], nullptr);
    const double G = std::strtod(argv[7], nullptr);
    r.out = argv[10];
    {
        std::stringstream ss(argv[11]);
        std::string tok;
        while (std::getline(ss, tok, ',')) r.dumps.push_back(std::atoi(tok.c_str()));
    }
    r.last = r.dumps.empty() ? 0 : *std::max_element(r.dumps.begin(), r.dumps.end());
    if (argv[11][0] == '+') {
        r.last = std::atoi(argv[11] + 1);
        r.dumps.clear();
    }
    double sol[2] = {0.0, 0.0};
    const bool film = argc == 14;
    if (film) {
        sol[0] = std::strtod(argv[12], nullptr);
        sol[1] = std::strtod(argv[13], nullptr);
    }
    r.p = slab_params(r.NX, r.NY, r.NZ);
    const size_t N = static_cast<size_t>(r.NX) * r.NY * r.NZ;
    r.rA = read_doubles(argv[8], N);
    r.rB = read_doubles(argv[9], N);
    const std::string inp =
        input_file(r.out, r.p.rank, q == 19 ? "D3Q19" : "D2Q9", tauA, tauB, G, film ? sol : nullptr);
    GetPot gp(inp.c_str());
    std::remove(inp.c_str());
    if (q == 19) {
        if (film) { mcmp3DFilm m(r.p, gp); run3d(m, r); }
        else { mcmp3D m(r.p, gp); run3d(m, r); }
    } else {
        if (film) { mcmp2DFilm m(r.p, gp); run2d(m, r); }
        else { mcmp2D m(r.p, gp); run2d(m, r); }
    }
    MPI::Finalize();
    return 0;
}
