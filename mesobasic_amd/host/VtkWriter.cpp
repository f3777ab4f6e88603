#include "VtkWriter.hpp"

#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace mesobasic {

namespace {

std::ofstream open_vtk(const std::string& dir, const std::string& tag, int step)
{
    std::stringstream name;
    name << dir << "/" << tag << "_" << step << ".vtk";
    std::ofstream out(name.str(), std::ios::out | std::ios::app);
    if (!out) throw std::runtime_error("cannot open " + name.str());
    return out;
}

}  // namespace

void write_vtk_3d(const std::string& dir, const std::string& tag, int step, int NX, int NY,
                  int NZ, int iskip, int jskip, int kskip, const std::vector<double>& rho)
{
    std::ofstream out = open_vtk(dir, tag, step);
    const std::string d = "   ";
    out << "# vtk DataFile Version 3.1" << std::endl;
    out << "VTK file containing grid data" << std::endl;
    out << "ASCII" << std::endl;
    out << " " << std::endl;
    out << "DATASET STRUCTURED_POINTS" << std::endl;
    out << "DIMENSIONS" << d << NX / iskip << d << NY / jskip << d << NZ / kskip << std::endl;
    out << "ORIGIN " << d << 0 << d << 0 << d << 0 << std::endl;
    out << "SPACING" << d << 1.0 * iskip << d << 1.0 * jskip << d << 1.0 * kskip << std::endl;
    out << " " << std::endl;
    out << "POINT_DATA " << (NX / iskip) * (NY / jskip) * (NZ / kskip) << std::endl;
    out << "SCALARS " << tag << " float" << std::endl;
    out << "LOOKUP_TABLE default" << std::endl;
    // loop order and selection of LBfluid3D.cpp:399-414 (one rank, xOffset 0)
    for (int k = 1; k < NZ + 1; k += kskip)
        for (int j = 1; j < NY + 1; j += jskip)
            for (int i = 1; i < NX + 1; i++) {
                if (i == 0 || i % iskip == 0) {
                    const long long c = (static_cast<long long>(i - 1) * NY + (j - 1)) * NZ + (k - 1);
                    out << std::fixed << std::setprecision(3) << rho[c] << std::endl;
                }
            }
}

void write_vtk_2d(const std::string& dir, const std::string& tag, int step, int NX, int NY,
                  int iskip, int jskip, const std::vector<double>& rho,
                  const std::vector<double>& ustar, const std::vector<double>& vstar)
{
    std::ofstream out = open_vtk(dir, tag, step);
    const std::string d = "   ";
    out << "# vtk DataFile Version 3.1" << std::endl;
    out << "VTK file containing grid data" << std::endl;
    out << "ASCII" << std::endl;
    out << " " << std::endl;
    out << "DATASET STRUCTURED_POINTS" << std::endl;
    out << "DIMENSIONS" << d << NX / iskip << d << NY / jskip << d << 1 << std::endl;
    out << "ORIGIN " << d << 1 << d << 1 << d << 0 << std::endl;
    out << "SPACING" << d << 1.0 * iskip << d << 1.0 * jskip << d << 1.0 << std::endl;
    out << " " << std::endl;
    out << "POINT_DATA " << (NX / iskip) * (NY / jskip) << std::endl;
    out << "SCALARS " << tag << " float" << std::endl;
    out << "LOOKUP_TABLE default" << std::endl;
    // LBfluid2D.cpp:440-453
    for (int j = 1; j < NY + 1; j += jskip)
        for (int i = 1; i < NX + 1; i++)
            if (i == 0 || i % iskip == 0)
                out << std::fixed << std::setprecision(3)
                    << rho[static_cast<long long>(i - 1) * NY + (j - 1)] << std::endl;
    // LBfluid2D.cpp:461-479
    out << "   " << std::endl;
    out << "VECTORS Velocity float" << std::endl;
    for (int j = 1; j < NY + 1; j += jskip)
        for (int i = 1; i < NX + 1; i++)
            if (i == 0 || i % iskip == 0) {
                const long long c = static_cast<long long>(i - 1) * NY + (j - 1);
                out << std::fixed << std::setprecision(3) << ustar[c] << " " << vstar[c] << " "
                    << 0.0 << std::endl;
            }
}

void write_model_vtk(const std::string& dir, const CommonParams& p, int q, int step,
                     const std::vector<double>& rho, const std::vector<double>& u,
                     const std::vector<double>& v, const std::vector<double>& fx,
                     const std::vector<double>& fy)
{
    if (q == 19) {
        write_vtk_3d(dir, "rhoA", step, p.NX, p.NY, p.NZ, p.iskip, p.jskip, p.kskip, rho);
        return;
    }
    std::vector<double> us(rho.size()), vs(rho.size());
    for (size_t c = 0; c < rho.size(); c++) {
        us[c] = u[c] + 0.5 * fx[c] / rho[c];
        vs[c] = v[c] + 0.5 * fy[c] / rho[c];
    }
    write_vtk_2d(dir, "rhoA", step, p.NX, p.NY, p.iskip, p.jskip, rho, us, vs);
}

void write_capsule_vtk(const std::string& dir, const std::string& tag, int step, const std::vector<double>& x,
                       const std::vector<double>& y)
{
    // IBcapsules2D::writeVTKFile, IBcapsules2D.cpp:492-562
    std::ofstream out = open_vtk(dir, tag, step);
    const size_t nn = x.size();
    out << "# vtk DataFile Version 3.1" << std::endl;
    out << "VTK file containing IBM data" << std::endl;
    out << "ASCII" << std::endl;
    out << " " << std::endl;
    out << "DATASET POLYDATA" << std::endl;
    out << " " << std::endl;
    out << "POINTS " << nn << " float" << std::endl;
    for (size_t n = 0; n < nn; n++)
        out << std::fixed << std::setprecision(3) << x[n] << "  " << y[n] << "  " << 0.0 << std::endl;
    out << " " << std::endl;
    out << "LINES " << nn << " " << 3 * nn << std::endl;
    for (size_t n = 0; n < nn; n++) out << "2  " << n << "  " << (n + 1 == nn ? 0 : n + 1) << std::endl;
    out << " " << std::endl;
    out << "VERTICES " << nn << " " << 2 * nn << std::endl;
    for (size_t n = 0; n < nn; n++) out << "1 " << n << std::endl;
}

}  // namespace mesobasic
