// ScmpModel.hpp -- single-component Shan-Chen model on the GPU, the common
// part of scmp3D (D3Q19) and scmp2D (D2Q9).  The lattice lives in HBM behind
// the C ABI (include/mesobasic_lbm.h); this class owns one handle.
//
// Keys (read as scmp3D.cpp:21-31 reads them, same defaults):
//   LBApp/G (-6.0)  LBApp/rhoAi (0.5)  LBApp/rhoAi_noise (0.1)
//   LBApp/tauA (1.0)  LBApp/stencil (D3Q19 for scmp3D, D2Q9 for scmp2D)
// Keys of this build only (the reference ignores unknown keys):
//   LBApp/seed       mt19937 seed of the noisy start (default: time(NULL),
//                    the reference's srand(time(NULL)) at rank 0)
//   LBApp/init       noise | droplet
//   LBApp/R, W, rho_l, rho_v   droplet radius, interface width, densities
//   LBApp/walls      1 = bounce-back walls on the last axis
//   LBApp/precision  fp64 | fp32 | fp16 | bf16
//   LBApp/rho_shift  shift of the 16-bit storage (default rhoAi)
//   LBApp/kernel     auto | two_pass | fused
//   LBApp/order      3 (scmp3D step order) | 2 (LBfluid2D step order)
//   Output/async     1 (default): VTK output overlapped with the next steps
//                    (OutputQueue); 0: written before the output returns
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/mesobasic_lbm.h"
#include "LBBaseClass.hpp"
#include "OutputQueue.hpp"

namespace mesobasic {

// The keys scmp3D reads (scmp3D.cpp:21-31) with the same defaults, the
// keys of this build listed above, and the initial density they give.
struct ScmpKeys {
    double G, rhoAi, rhoAi_noise, tauA;
    std::string init_kind;
    unsigned seed;
    double R, W, rho_l, rho_v;

    static ScmpKeys read(const CommonParams& p, const InputParams& in, int q);
    // rho0 of the lattice, interior order (scmp3D.cpp:63-76 with a seeded
    // mt19937 in place of rand(), or the droplet of LBApp/init = droplet)
    std::vector<double> initial_rho(const CommonParams& p, int q) const;
};

class ScmpModel : public LBBaseClass {
public:
    ScmpModel(const CommonParams& p, const InputParams& in, int q, int default_order);
    ~ScmpModel() override;
    ScmpModel(const ScmpModel&) = delete;
    ScmpModel& operator=(const ScmpModel&) = delete;

    void setTimeStep(int step) override { current_step = step; }
    void initLatticeBoltzmann() override;
    void updateLatticeBoltzmann() override;
    // scmp3D.cpp:127-133 (density) for D3Q19; for D2Q9 density and the
    // Shan-Chen velocity as mcmp2D.cpp:131-137
    void outputLatticeBoltzmann() override;
    void finishOutput() override;

    // Host copies of the reference's r,u,v,w arrays (interior, x slowest).
    void fields(std::vector<double>& rho, std::vector<double>& u, std::vector<double>& v,
                std::vector<double>& w);
    void force(std::vector<double>& fx, std::vector<double>& fy, std::vector<double>& fz);
    lbm_handle* handle() { return h_; }
    long long nodes() const { return static_cast<long long>(p.nx) * p.ny * p.nz; }

protected:
    const CommonParams& p;
    int q_;
    int current_step = 0;
    ScmpKeys k_;
    lbm_handle* h_ = nullptr;
    std::unique_ptr<OutputQueue> out_;

    void check(int rc, const char* what) const;
};

class scmp3D : public ScmpModel {
public:
    scmp3D(const CommonParams& p, const InputParams& in) : ScmpModel(p, in, 19, 3) {}
};

// D2Q9 single-component Shan-Chen: the scmp3D force on the D2Q9 stencil in
// the LBfluid2D step order (LBfluid2D.cpp:227-312).  The reference has no such
// class; its factory type name here is "scmp2D".
class scmp2D : public ScmpModel {
public:
    scmp2D(const CommonParams& p, const InputParams& in) : ScmpModel(p, in, 9, 2) {}
};

// Two-component Shan-Chen: mcmp2D (D2Q9, mcmp2D.cpp) and mcmp3D (D3Q19,
// mcmp3D.cpp).  One handle holds both components (lbm_desc.components = 2).
//
// Keys (read as mcmp2D.cpp:20-34 / mcmp3D.cpp:20-34 read them, same
// defaults): LBApp/G (3.0), rhoAi and rhoBi (0.5), rhoAi_noise and
// rhoBi_noise (0.1), tauA and tauB (1.0), stencil.  Keys of this build only:
// LBApp/seed (mt19937 seed; default time(NULL)), LBApp/init (noise |
// droplet: an A-rich drop of radius R, width W, densities rho_l / rho_v in
// B, with rhoB = rho_l + rho_v - rhoA).
//
// film = true: mcmp2DFilm / mcmp3DFilm (walls on the last axis; keys
// LBApp/rhoA_sol and rhoB_sol, default 0.5, mcmp2DFilm.cpp:21-22).  As in the
// reference, mcmp3DFilm's start fills the fluid rows only and leaves the wall
// rows at density 0 (mcmp3DFilm.cpp:71-84).
class McmpModel : public LBBaseClass {
public:
    McmpModel(const CommonParams& p, const InputParams& in, int q, bool film = false);
    ~McmpModel() override;
    McmpModel(const McmpModel&) = delete;
    McmpModel& operator=(const McmpModel&) = delete;

    void setTimeStep(int step) override { current_step = step; }
    void initLatticeBoltzmann() override;
    void updateLatticeBoltzmann() override;
    void outputLatticeBoltzmann() override;
    void finishOutput() override;

    // Host copies of component comp's r,u,v,w arrays and force arrays.
    void fields(int comp, std::vector<double>& rho, std::vector<double>& u, std::vector<double>& v,
                std::vector<double>& w);
    void force(int comp, std::vector<double>& fx, std::vector<double>& fy, std::vector<double>& fz);
    lbm_handle* handle() { return h_; }
    long long nodes() const { return static_cast<long long>(p.nx) * p.ny * p.nz; }

private:
    const CommonParams& p;
    int q_;
    int current_step = 0;
    bool film_;
    double G, rhoAi, rhoBi, rhoAi_noise, rhoBi_noise, tauA, tauB, rhoA_sol, rhoB_sol;
    std::string init_kind;
    unsigned seed;
    double R, W, rho_l, rho_v;
    lbm_handle* h_ = nullptr;
    std::unique_ptr<OutputQueue> out_;

    void check(int rc, const char* what) const;
};

// Immersed-boundary capsule in a channel (capsule2D, capsule2D.cpp:10-166
// with IBcapsules2D): a D2Q9 fluid in the LBfluid2D step order with
// bounce-back walls in y, driven by a body force on the rows between the
// walls, and a circular capsule of nnodes IB nodes tied to their rest
// positions.  Keys (read as capsule2D.cpp:25-31 and IBcapsules2D.cpp:45-48
// read them, same defaults): LBApp/rhoAi (0.5), rhoAi_noise (0.1), tauA
// (1.0), stencil (D2Q9), ncaps (1; one capsule is placed, as in the
// reference), nnodes (20), h (2.0; the footprint half-width of the cosine
// kernel, which the bilinear kernel the reference uses ignores), kstiff
// (1.0); Time/dt (1.0).  Keys of this build only: LBApp/seed, LBApp/capsule_x,
// capsule_y, capsule_r (the reference's fixed circle 30, 30, 8:
// capsule2D.cpp:113), body_fx, body_fy (its fixed 0.000008, 0:
// capsule2D.cpp:132-133).
class capsule2D : public LBBaseClass {
public:
    capsule2D(const CommonParams& p, const InputParams& in);
    ~capsule2D() override;
    capsule2D(const capsule2D&) = delete;
    capsule2D& operator=(const capsule2D&) = delete;

    void setTimeStep(int step) override { current_step = step; }
    void initLatticeBoltzmann() override;
    void updateLatticeBoltzmann() override;
    void outputLatticeBoltzmann() override;
    void finishOutput() override;
    lbm_handle* handle() { return h_; }

private:
    const CommonParams& p;
    int current_step = 0;
    double rhoAi, rhoAi_noise, tauA;
    int nnodes;
    double kstiff, cx, cy, cr;
    unsigned seed;
    lbm_handle* h_ = nullptr;
    std::unique_ptr<OutputQueue> out_;

    void check(int rc, const char* what) const;
};

class mcmp2D : public McmpModel {
public:
    mcmp2D(const CommonParams& p, const InputParams& in) : McmpModel(p, in, 9) {}
};

class mcmp3D : public McmpModel {
public:
    mcmp3D(const CommonParams& p, const InputParams& in) : McmpModel(p, in, 19) {}
};

class mcmp2DFilm : public McmpModel {
public:
    mcmp2DFilm(const CommonParams& p, const InputParams& in) : McmpModel(p, in, 9, true) {}
};

class mcmp3DFilm : public McmpModel {
public:
    mcmp3DFilm(const CommonParams& p, const InputParams& in) : McmpModel(p, in, 19, true) {}
};

}  // namespace mesobasic
