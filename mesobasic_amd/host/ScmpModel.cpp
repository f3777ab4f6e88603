#include "ScmpModel.hpp"

#include <cmath>
#include <ctime>
#include <random>
#include <stdexcept>
#include <string>

#include "VtkWriter.hpp"

namespace mesobasic {

namespace {

int precision_of(const std::string& s)
{
    if (s == "fp64") return LBM_FP64;
    if (s == "fp32") return LBM_FP32;
    if (s == "fp16") return LBM_FP16S;
    if (s == "bf16") return LBM_BF16S;
    if (s == "fp16n") return LBM_FP16N;
    if (s == "bf16n") return LBM_BF16N;
    throw std::runtime_error("LBApp/precision must be fp64, fp32, fp16, bf16, fp16n or bf16n (got " + s + ")");
}

int kernel_of(const std::string& s)
{
    if (s == "auto") return LBM_KERNEL_AUTO;
    if (s == "two_pass") return LBM_KERNEL_TWO_PASS;
    if (s == "fused") return LBM_KERNEL_FUSED;
    throw std::runtime_error("LBApp/kernel must be auto, two_pass or fused (got " + s + ")");
}

}  // namespace

ScmpKeys ScmpKeys::read(const CommonParams& p, const InputParams& in, int q)
{
    // scmp3D.cpp:21-31
    ScmpKeys k;
    k.G = in("LBApp/G", -6.0);
    k.rhoAi = in("LBApp/rhoAi", 0.5);
    k.rhoAi_noise = in("LBApp/rhoAi_noise", 0.1);
    k.tauA = in("LBApp/tauA", 1.0);
    const std::string stype = in("LBApp/stencil", q == 19 ? "D3Q19" : "D2Q9");
    if ((q == 19 && stype != "D3Q19") || (q == 9 && stype != "D2Q9"))
        throw std::runtime_error("LBApp/stencil " + stype + " does not match this model");
    k.init_kind = in("LBApp/init", "noise");
    if (k.init_kind != "noise" && k.init_kind != "droplet")
        throw std::runtime_error("LBApp/init must be noise or droplet (got " + k.init_kind + ")");
    k.seed = in.has("LBApp/seed") ? static_cast<unsigned>(in("LBApp/seed", 0.0))
                                  : static_cast<unsigned>(std::time(nullptr) * (p.rank + 1));
    k.R = in("LBApp/R", 0.25 * p.NX);
    k.W = in("LBApp/W", 2.0);
    k.rho_l = in("LBApp/rho_l", 2.0);
    k.rho_v = in("LBApp/rho_v", 0.15);
    if (q == 9 && p.nz != 1) throw std::runtime_error("scmp2D needs Domain/nz = 1");
    return k;
}

std::vector<double> ScmpKeys::initial_rho(const CommonParams& p, int q) const
{
    const int nz = (q == 19) ? p.nz : 1;
    std::vector<double> rho(static_cast<size_t>(p.nx) * p.ny * nz);
    int rc = (init_kind == "droplet")
                 ? lbm_fill_droplet(p.nx, p.ny, nz, R, W, rho_l, rho_v, rho.data())
                 : lbm_fill_noise(seed, static_cast<int64_t>(rho.size()), rhoAi, rhoAi_noise, rho.data());
    if (rc != LBM_OK) throw std::runtime_error(std::string("initial density: ") + lbm_last_error());
    return rho;
}

ScmpModel::ScmpModel(const CommonParams& pin, const InputParams& in, int q, int default_order)
    : p(pin), q_(q), k_(ScmpKeys::read(pin, in, q))
{
    lbm_desc d{};
    d.q = q;
    d.nx = p.nx;
    d.ny = p.ny;
    d.nz = (q == 19) ? p.nz : 1;
    d.tau = k_.tauA;
    d.G = k_.G;
    d.order = in("LBApp/order", default_order);
    d.walls = in("LBApp/walls", 0);
    d.precision = precision_of(in("LBApp/precision", "fp64"));
    d.rho_shift = in("LBApp/rho_shift", k_.rhoAi);
    d.kernel = kernel_of(in("LBApp/kernel", "auto"));
    check(lbm_create(&d, &h_), "lbm_create");
    out_.reset(new OutputQueue(p, q, in("Output/async", 1) != 0));
}

ScmpModel::~ScmpModel() { lbm_destroy(h_); }

void ScmpModel::check(int rc, const char* what) const
{
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

void ScmpModel::initLatticeBoltzmann()
{
    // scmp3D.cpp:54-82: initial density, zero velocity, ghost rho, force,
    // f = feq
    const std::vector<double> rho = k_.initial_rho(p, q_);
    check(lbm_set_rho(h_, rho.data()), "lbm_set_rho");
    check(lbm_init_equilibrium(h_), "lbm_init_equilibrium");
}

void ScmpModel::updateLatticeBoltzmann()
{
    // scmp3D.cpp:90-119 (the Barrier of :97 has no counterpart on one GPU)
    check(lbm_step(h_, 1), "lbm_step");
}

void ScmpModel::fields(std::vector<double>& rho, std::vector<double>& u, std::vector<double>& v,
                       std::vector<double>& w)
{
    const size_t n = static_cast<size_t>(nodes());
    rho.resize(n); u.resize(n); v.resize(n); w.resize(n);
    check(lbm_get_fields(h_, LBM_FIELDS_REFERENCE, rho.data(), u.data(), v.data(), w.data()),
          "lbm_get_fields");
}

void ScmpModel::force(std::vector<double>& fx, std::vector<double>& fy, std::vector<double>& fz)
{
    const size_t n = static_cast<size_t>(nodes());
    fx.resize(n); fy.resize(n); fz.resize(n);
    check(lbm_get_force(h_, fx.data(), fy.data(), fz.data()), "lbm_get_force");
}

void ScmpModel::outputLatticeBoltzmann() { out_->output(h_, 0, current_step); }

void ScmpModel::finishOutput() { out_->finish(h_); }

McmpModel::McmpModel(const CommonParams& pin, const InputParams& in, int q, bool film)
    : p(pin), q_(q), film_(film)
{
    // mcmp2D.cpp:18-37 / mcmp3D.cpp:18-37 (the film forms add the wall
    // densities, mcmp2DFilm.cpp:21-22, mcmp3DFilm.cpp:21-22)
    G = in("LBApp/G", 3.0);
    rhoA_sol = in("LBApp/rhoA_sol", 0.5);
    rhoB_sol = in("LBApp/rhoB_sol", 0.5);
    rhoAi = in("LBApp/rhoAi", 0.5);
    rhoBi = in("LBApp/rhoBi", 0.5);
    rhoAi_noise = in("LBApp/rhoAi_noise", 0.1);
    rhoBi_noise = in("LBApp/rhoBi_noise", 0.1);
    tauA = in("LBApp/tauA", 1.0);
    tauB = in("LBApp/tauB", 1.0);
    const std::string stype = in("LBApp/stencil", q == 19 ? "D3Q19" : "D2Q9");
    if ((q == 19 && stype != "D3Q19") || (q == 9 && stype != "D2Q9"))
        throw std::runtime_error("LBApp/stencil " + stype + " does not match this model");
    init_kind = in("LBApp/init", "noise");
    seed = in.has("LBApp/seed") ? static_cast<unsigned>(in("LBApp/seed", 0.0))
                                : static_cast<unsigned>(std::time(nullptr) * (p.rank + 1));
    R = in("LBApp/R", 0.25 * p.NX);
    W = in("LBApp/W", 2.0);
    rho_l = in("LBApp/rho_l", 1.0);
    rho_v = in("LBApp/rho_v", 0.05);
    if (q == 9 && p.nz != 1) throw std::runtime_error("mcmp2D needs Domain/nz = 1");

    lbm_desc d{};
    d.q = q;
    d.nx = p.nx;
    d.ny = p.ny;
    d.nz = (q == 19) ? p.nz : 1;
    d.tau = tauA;
    d.G = G;
    d.order = (q == 19) ? 3 : 2;   // mcmp3D.cpp:97-123 / mcmp2D.cpp:97-123
    d.precision = LBM_FP64;
    d.kernel = LBM_KERNEL_AUTO;
    d.components = 2;
    d.tau_b = tauB;
    d.walls = film ? 1 : 0;
    d.rho_sol_a = rhoA_sol;
    d.rho_sol_b = rhoB_sol;
    check(lbm_create(&d, &h_), "lbm_create");
    out_.reset(new OutputQueue(p, q, in("Output/async", 1) != 0));
}

McmpModel::~McmpModel() { lbm_destroy(h_); }

void McmpModel::check(int rc, const char* what) const
{
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

void McmpModel::initLatticeBoltzmann()
{
    // mcmp2D.cpp:58-89 / mcmp3D.cpp:58-93: per node in loop order, one draw
    // for A then one for B; a seeded mt19937 in place of rand()/RAND_MAX
    const size_t n = static_cast<size_t>(nodes());
    std::vector<double> rA(n), rB(n);
    if (init_kind == "noise") {
        // mcmp3DFilm fills the fluid rows only (k = 2 .. nz-1); its wall rows
        // keep density 0 (mcmp3DFilm.cpp:71-84)
        const bool skip_walls = film_ && q_ == 19;
        std::mt19937 gen(seed);
        for (size_t c = 0; c < n; c++) {
            const int z = static_cast<int>(c % static_cast<size_t>(q_ == 19 ? p.nz : 1));
            if (skip_walls && (z == 0 || z == p.nz - 1)) {
                rA[c] = 0.0;
                rB[c] = 0.0;
                continue;
            }
            const double xa = static_cast<double>(gen()) / 4294967295.0;
            const double xb = static_cast<double>(gen()) / 4294967295.0;
            rA[c] = rhoAi + rhoAi_noise * (xa - 0.5);
            rB[c] = rhoBi + rhoBi_noise * (xb - 0.5);
        }
    } else if (init_kind == "droplet") {
        check(lbm_fill_droplet(p.nx, p.ny, q_ == 19 ? p.nz : 1, R, W, rho_l, rho_v, rA.data()),
              "lbm_fill_droplet");
        for (size_t c = 0; c < n; c++) rB[c] = (rho_l + rho_v) - rA[c];
    } else {
        throw std::runtime_error("LBApp/init must be noise or droplet (got " + init_kind + ")");
    }
    check(lbm_set_rho_component(h_, 0, rA.data()), "lbm_set_rho_component");
    check(lbm_set_rho_component(h_, 1, rB.data()), "lbm_set_rho_component");
    check(lbm_init_equilibrium(h_), "lbm_init_equilibrium");
}

void McmpModel::updateLatticeBoltzmann()
{
    // mcmp2D.cpp:97-123 / mcmp3D.cpp:97-123 (the Barrier has no counterpart
    // on one GPU)
    check(lbm_step(h_, 1), "lbm_step");
}

void McmpModel::fields(int comp, std::vector<double>& rho, std::vector<double>& u,
                       std::vector<double>& v, std::vector<double>& w)
{
    const size_t n = static_cast<size_t>(nodes());
    rho.resize(n); u.resize(n); v.resize(n); w.resize(n);
    check(lbm_get_fields_component(h_, comp, LBM_FIELDS_REFERENCE, rho.data(), u.data(), v.data(),
                                   w.data()),
          "lbm_get_fields_component");
}

void McmpModel::force(int comp, std::vector<double>& fx, std::vector<double>& fy,
                      std::vector<double>& fz)
{
    const size_t n = static_cast<size_t>(nodes());
    fx.resize(n); fy.resize(n); fz.resize(n);
    check(lbm_get_force_component(h_, comp, fx.data(), fy.data(), fz.data()), "lbm_get_force_component");
}

// mcmp2D.cpp:131-137 / mcmp3D.cpp:131-138: component A only
void McmpModel::outputLatticeBoltzmann() { out_->output(h_, 0, current_step); }

void McmpModel::finishOutput() { out_->finish(h_); }

capsule2D::capsule2D(const CommonParams& pin, const InputParams& in) : p(pin)
{
    // capsule2D.cpp:25-37, IBcapsules2D.cpp:45-48
    rhoAi = in("LBApp/rhoAi", 0.5);
    rhoAi_noise = in("LBApp/rhoAi_noise", 0.1);
    tauA = in("LBApp/tauA", 1.0);
    const std::string stype = in("LBApp/stencil", "D2Q9");
    if (stype != "D2Q9") throw std::runtime_error("LBApp/stencil " + stype + " does not match capsule2D");
    if (p.nz != 1) throw std::runtime_error("capsule2D needs Domain/nz = 1");
    nnodes = in("LBApp/nnodes", 20);
    kstiff = in("LBApp/kstiff", 1.0);
    cx = in("LBApp/capsule_x", 30.0);
    cy = in("LBApp/capsule_y", 30.0);
    cr = in("LBApp/capsule_r", 8.0);
    seed = in.has("LBApp/seed") ? static_cast<unsigned>(in("LBApp/seed", 0.0))
                                : static_cast<unsigned>(std::time(nullptr) * (p.rank + 1));
    lbm_desc d{};
    d.q = 9;
    d.nx = p.nx;
    d.ny = p.ny;
    d.nz = 1;
    d.tau = tauA;
    d.G = 0.0;
    d.order = 2;   // LBfluid2D::updateFluid_SC (capsule2D.cpp:142-145)
    d.walls = 1;   // yBC = 1
    d.precision = LBM_FP64;
    d.kernel = LBM_KERNEL_TWO_PASS;
    check(lbm_create(&d, &h_), "lbm_create");
    lbm_capsule_desc cd{};
    cd.nnodes = nnodes;
    cd.kstiff = kstiff;
    cd.dt = p.dt;
    cd.body_fx = in("LBApp/body_fx", 0.000008);
    cd.body_fy = in("LBApp/body_fy", 0.0);
    check(lbm_capsule_attach(h_, &cd), "lbm_capsule_attach");
    out_.reset(new OutputQueue(p, 9, in("Output/async", 1) != 0, true));
}

capsule2D::~capsule2D() { lbm_destroy(h_); }

void capsule2D::check(int rc, const char* what) const
{
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

void capsule2D::initLatticeBoltzmann()
{
    // capsule2D.cpp:84-120: noisy density on every node (a seeded mt19937 in
    // place of rand()), zero velocity and forces, the circular capsule
    // (IBcapsules2D::initCircularCapsule, IBcapsules2D.cpp:130-141), f = feq
    std::vector<double> rho(static_cast<size_t>(p.nx) * p.ny);
    check(lbm_fill_noise(seed, static_cast<int64_t>(rho.size()), rhoAi, rhoAi_noise, rho.data()),
          "lbm_fill_noise");
    std::vector<double> x(static_cast<size_t>(nnodes)), y(static_cast<size_t>(nnodes));
    double theta = 0.0;
    const double dTheta = (2.0 * 3.14159265) / double(nnodes);
    for (int n = 0; n < nnodes; n++) {
        x[n] = cx + cr * std::cos(theta);
        y[n] = cy + cr * std::sin(theta);
        theta += dTheta;
    }
    check(lbm_capsule_set_nodes(h_, x.data(), y.data()), "lbm_capsule_set_nodes");
    check(lbm_set_rho(h_, rho.data()), "lbm_set_rho");
    check(lbm_init_equilibrium(h_), "lbm_init_equilibrium");
}

void capsule2D::updateLatticeBoltzmann()
{
    // capsule2D.cpp:123-150 (the Barrier has no counterpart on one GPU)
    check(lbm_step(h_, 1), "lbm_step");
}

// capsule2D.cpp:158-164: the fluid (density and u*) and the IB mesh
void capsule2D::outputLatticeBoltzmann() { out_->output(h_, 0, current_step); }

void capsule2D::finishOutput() { out_->finish(h_); }

// LBBaseClass::LBFactory (LBBaseClass.cpp:21-45), with the reference's
// default type mcmp2D; unknown names yield nullptr.
LBBaseClass* LBBaseClass::LBFactory(const CommonParams& p, const InputParams& in)
{
    const std::string lb_type = in("LBApp/type", "mcmp2D");
    if (lb_type == "scmp3D") return new scmp3D(p, in);
    if (lb_type == "scmp2D") return new scmp2D(p, in);
    if (lb_type == "mcmp2D") return new mcmp2D(p, in);
    if (lb_type == "mcmp3D") return new mcmp3D(p, in);
    if (lb_type == "mcmp2DFilm") return new mcmp2DFilm(p, in);
    if (lb_type == "mcmp3DFilm") return new mcmp3DFilm(p, in);
    if (lb_type == "capsule2D") return new capsule2D(p, in);
    return nullptr;
}

}  // namespace mesobasic
