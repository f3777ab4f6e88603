// ref_harness.cpp -- drives the reference's own lattice-Boltzmann classes
// (compiled in place from the reference tree by oracle/Makefile.ref; no
// reference source is copied) with an injected initial density, and dumps
// full-precision binary fields.  TEST INFRASTRUCTURE ONLY.
//
// Why a harness: the shipped model seeds rand() from time() (scmp3D.cpp:61)
// and its VTK writer keeps 3 decimals (LBfluid3D.cpp:407), so neither the
// initial state nor the output can carry a 1e-12 comparison.
//
// D3Q19: the reference scmp3D class itself (scmp3D.cpp:10-190), stepped as
// LBApp::stepForward does (LBApp.cpp:98-114); z walls by calling
// LBfluid3D::bounceBackWallsZdir after each update, the order
// mcmp3DFilm.cpp:132-140 uses.
// D2Q9: the reference has no single-component 2-D class, so the harness
// composes LBfluid2D (LBfluid2D.cpp:201-312, used verbatim) with the scmp3D
// force formula (scmp3D.cpp:141-179) evaluated on the D2Q9 stencil, in the
// step order of mcmp2D::updateLatticeBoltzmann (mcmp2D.cpp:97-123).
//
// Usage (one line per option):
//   ref_harness q nx ny nz tau G walls rho0.bin out_prefix step[,step...]
// Writes out_prefix_<step>.bin = rho | u | v | w | f  (float64, interior
// nodes, x slowest and the last axis fastest; f is [node][q]).  In 2-D the w
// block is zeros.  Run under mpirun -np N to split x into slabs
// (LBApp.cpp:46-60); results do not depend on N.

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdarg>
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
#include "lattice_boltzmann/LBTypes/scmp3D.hpp"
#include "lattice_boltzmann/LBUtils/LBfluid2D.hpp"
#undef private

namespace {

// Slab split of LBApp::LBApp (LBApp.cpp:29-60), restated for the harness.
CommonParams make_params(int NX, int NY, int NZ)
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
    int nxAve = NX / p.np, nxExtra = NX % p.np;
    p.nx = ((p.rank + 1) <= nxExtra) ? nxAve + 1 : nxAve;
    int xOff = 0;
    for (int i = 0; i < p.rank; i++) xOff += (i + 1 <= nxExtra) ? nxAve + 1 : nxAve;
    p.ny = NY; p.nz = NZ; p.xOff = xOff;
    p.LX = NX; p.LY = NY; p.LZ = NZ;
    return p;
}

std::vector<double> read_all(const char* path, size_t n)
{
    std::vector<double> v(n);
    FILE* fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(double), n, fp) != n) {
        std::fprintf(stderr, "ref_harness: cannot read %zu doubles from %s\n", n, path);
        MPI::COMM_WORLD.Abort(2);
    }
    std::fclose(fp);
    return v;
}

// Gathers rank slabs (x is slowest, so the global array is the rank-ordered
// concatenation) and writes them on rank 0.
void dump(const std::string& path, const std::vector<double>& local_blocks,
          int per_node_blocks_total, long local_nodes)
{
    int np = MPI::COMM_WORLD.Get_size(), rank = MPI::COMM_WORLD.Get_rank();
    int cnt = (int)local_blocks.size();
    std::vector<int> counts(np), displs(np);
    MPI::COMM_WORLD.Gather(&cnt, 1, MPI::INT, counts.data(), 1, MPI::INT, 0);
    std::vector<double> all;
    if (rank == 0) {
        int tot = 0;
        for (int r = 0; r < np; r++) { displs[r] = tot; tot += counts[r]; }
        all.resize(tot);
    }
    MPI::COMM_WORLD.Gatherv(local_blocks.data(), cnt, MPI::DOUBLE, all.data(),
                            counts.data(), displs.data(), MPI::DOUBLE, 0);
    if (rank != 0) return;
    // Each rank packed [rho|u|v|w|f] for its slab; reorder to field-major.
    (void)per_node_blocks_total; (void)local_nodes;
    std::vector<long> rank_nodes(np);
    for (int r = 0; r < np; r++) rank_nodes[r] = counts[r];
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) MPI::COMM_WORLD.Abort(3);
    // fields 0..3 are scalars (1 per node), field 4 is f (q per node)
    for (int field = 0; field < 5; field++) {
        for (int r = 0; r < np; r++) {
            const double* base = all.data() + displs[r];
            long n_per = counts[r] / (4 + per_node_blocks_total);  // nodes
            long off = (field < 4) ? field * n_per : 4 * n_per;
            long len = (field < 4) ? n_per : n_per * per_node_blocks_total;
            std::fwrite(base + off, sizeof(double), len, fp);
        }
    }
    std::fclose(fp);
}

// scmp3D formula (scmp3D.cpp:141-179) on the D2Q9 stencil over an LBfluid2D.
void forces_2d(LBfluid2D& fA, const Stencil& s, const CommonParams& p, double G)
{
    int deli = p.ny + 2, delj = 1;
    for (int i = 1; i < p.nx + 1; i++) {
        for (int j = 1; j < p.ny + 1; j++) {
            int ndx = i * deli + j * delj;
            double rhoA = fA.getRho(ndx);
            double psi_rhoA = (1.0 - exp(-rhoA));
            double GsumxA = 0.0, GsumyA = 0.0;
            for (int n = 0; n < s.nn; n++) {
                int nbr = (i + s.exi[n]) * deli + (j + s.eyi[n]) * delj;
                double psi_nbr = (1.0 - exp(-fA.getRho(nbr)));
                GsumxA += s.wa[n] * s.ex[n] * psi_nbr;
                GsumyA += s.wa[n] * s.ey[n] * psi_nbr;
            }
            fA.setFx(ndx, -G * psi_rhoA * GsumxA);
            fA.setFy(ndx, -G * psi_rhoA * GsumyA);
        }
    }
}

}  // namespace

int main(int argc, char** argv)
{
    MPI::Init();
    if (argc != 11) {
        if (MPI::COMM_WORLD.Get_rank() == 0)
            std::fprintf(stderr, "usage: ref_harness q nx ny nz tau G walls rho0.bin out_prefix steps\n");
        MPI::Finalize();
        return 2;
    }
    int q = std::atoi(argv[1]);
    int NX = std::atoi(argv[2]), NY = std::atoi(argv[3]), NZ = std::atoi(argv[4]);
    double tau = std::strtod(argv[5], nullptr), G = std::strtod(argv[6], nullptr);
    int walls = std::atoi(argv[7]);
    std::string out = argv[9];
    std::vector<int> dumps;
    {
        std::stringstream ss(argv[10]);
        std::string tok;
        while (std::getline(ss, tok, ',')) dumps.push_back(std::atoi(tok.c_str()));
    }
    int last = dumps.empty() ? 0 : *std::max_element(dumps.begin(), dumps.end());
    CommonParams p = make_params(NX, NY, NZ);
    int rank = p.rank;
    std::vector<double> rho0 = read_all(argv[8], (size_t)NX * NY * NZ);

    if (q == 19) {
        // GetPot keys read by scmp3D's constructor (scmp3D.cpp:21-31).
        char inp[256];
        std::snprintf(inp, sizeof(inp), "%s_rank%d_input.dat", out.c_str(), rank);
        {
            std::ofstream o(inp);
            o << std::setprecision(17);
            o << "[LBApp]\n  type = scmp3D\n  G = " << G << "\n  tauA = " << tau
              << "\n  stencil = D3Q19\n[]\n";
        }
        GetPot gp(inp);
        std::remove(inp);
        scmp3D m(p, gp);
        // initLatticeBoltzmann (scmp3D.cpp:54-82) with rho0 injected.
        for (int i = 1; i < p.nx + 1; i++)
            for (int j = 1; j < p.ny + 1; j++)
                for (int k = 1; k < p.nz + 1; k++) {
                    int ndx = i * m.deli + j * m.delj + k * m.delk;
                    long g = ((long)(i - 1 + p.xOff) * NY + (j - 1)) * NZ + (k - 1);
                    m.fA.setRho(ndx, rho0[g]);
                    m.fA.setU(ndx, 0.0); m.fA.setV(ndx, 0.0); m.fA.setW(ndx, 0.0);
                }
        m.fA.ghostNodesRho();
        m.calculateShanChenForces();
        m.fA.setFtoFeq(m.s);
        for (int step = 0; step <= last; step++) {
            if (step > 0) {
                m.setTimeStep(step);
                m.updateLatticeBoltzmann();
                if (walls) m.fA.bounceBackWallsZdir(m.s);
            }
            if (std::find(dumps.begin(), dumps.end(), step) == dumps.end()) continue;
            long nloc = (long)p.nx * p.ny * p.nz;
            std::vector<double> blk((4 + 19) * nloc);
            long c = 0;
            for (int i = 1; i < p.nx + 1; i++)
                for (int j = 1; j < p.ny + 1; j++)
                    for (int k = 1; k < p.nz + 1; k++, c++) {
                        int ndx = i * m.deli + j * m.delj + k * m.delk;
                        blk[0 * nloc + c] = m.fA.r[ndx];
                        blk[1 * nloc + c] = m.fA.u[ndx];
                        blk[2 * nloc + c] = m.fA.v[ndx];
                        blk[3 * nloc + c] = m.fA.w[ndx];
                        for (int n = 0; n < 19; n++)
                            blk[4 * nloc + c * 19 + n] = m.fA.f[(long)ndx * 19 + n];
                    }
            dump(out + "_" + std::to_string(step) + ".bin", blk, 19, nloc);
        }
    } else {
        LBfluid2D fA(p);
        Stencil s;
        s.setStencil("D2Q9");
        fA.setTau(tau);
        fA.allocateFs(s);
        int deli = p.ny + 2, delj = 1;
        for (int i = 1; i < p.nx + 1; i++)
            for (int j = 1; j < p.ny + 1; j++) {
                int ndx = i * deli + j * delj;
                long g = (long)(i - 1 + p.xOff) * NY + (j - 1);
                fA.setRho(ndx, rho0[g]);
                fA.setU(ndx, 0.0); fA.setV(ndx, 0.0);
            }
        fA.ghostNodesRho();
        forces_2d(fA, s, p, G);
        fA.setFtoFeq(s);
        for (int step = 0; step <= last; step++) {
            if (step > 0) {
                MPI::COMM_WORLD.Barrier();
                forces_2d(fA, s, p, G);
                fA.updateFluid_SC(s, 0, walls ? 1 : 0, true);
            }
            if (std::find(dumps.begin(), dumps.end(), step) == dumps.end()) continue;
            long nloc = (long)p.nx * p.ny;
            std::vector<double> blk((4 + 9) * nloc);
            long c = 0;
            for (int i = 1; i < p.nx + 1; i++)
                for (int j = 1; j < p.ny + 1; j++, c++) {
                    int ndx = i * deli + j * delj;
                    blk[0 * nloc + c] = fA.r[ndx];
                    blk[1 * nloc + c] = fA.u[ndx];
                    blk[2 * nloc + c] = fA.v[ndx];
                    blk[3 * nloc + c] = 0.0;
                    for (int n = 0; n < 9; n++)
                        blk[4 * nloc + c * 9 + n] = fA.f[(long)ndx * 9 + n];
                }
            dump(out + "_" + std::to_string(step) + ".bin", blk, 9, nloc);
        }
    }
    MPI::Finalize();
    return 0;
}
