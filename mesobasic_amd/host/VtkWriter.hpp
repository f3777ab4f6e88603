// VtkWriter.hpp -- legacy ASCII VTK files in the reference's exact format:
// LBfluid3D::writeVTKFile (LBfluid3D.cpp:355-422) for 3-D density and
// LBfluid2D::writeVTKFile (LBfluid2D.cpp:396-487) for 2-D density plus the
// Shan-Chen velocity.  Values are printed fixed with 3 decimals, x fastest.
#pragma once

#include <string>
#include <vector>

#include "CommonParams.h"

namespace mesobasic {

// Longest text printf("%.3f") gives for a double: sign, 309 integer
// digits, the point and 3 decimals.
constexpr int kMaxFixed3 = 320;

// Writes v at p as printf("%.3f") would in the C locale and returns the end
// (no terminating zero); p must have kMaxFixed3 bytes of room.
char* put_fixed3(char* p, double v);

// rho: interior array (x slowest, z fastest) of NX*NY*NZ values.
void write_vtk_3d(const std::string& dir, const std::string& tag, int step, int NX, int NY,
                  int NZ, int iskip, int jskip, int kskip, const std::vector<double>& rho);

// rho, ustar, vstar: interior arrays (x slowest) of NX*NY values.
void write_vtk_2d(const std::string& dir, const std::string& tag, int step, int NX, int NY,
                  int iskip, int jskip, const std::vector<double>& rho,
                  const std::vector<double>& ustar, const std::vector<double>& vstar);

// The IB mesh of a capsule (IBcapsules2D::writeVTKFile, IBcapsules2D.cpp:
// 492-562): node positions with 3 decimals, the closed polyline through them
// and one vertex per node.  Written as dir/<tag>_<step>.vtk.
void write_capsule_vtk(const std::string& dir, const std::string& tag, int step, const std::vector<double>& x,
                       const std::vector<double>& y);

// The output file of one step of a model: for D3Q19 (q = 19) the density
// (scmp3D.cpp:127-133); for D2Q9 the density and the Shan-Chen velocity
// u + F/(2 rho) (mcmp2D.cpp:131-137, LBfluid2D::getUStar, LBfluid2D.cpp:
// 164-172).  Written as dir/rhoA_<step>.vtk.
void write_model_vtk(const std::string& dir, const CommonParams& p, int q, int step,
                     const std::vector<double>& rho, const std::vector<double>& u,
                     const std::vector<double>& v, const std::vector<double>& fx,
                     const std::vector<double>& fy);

}  // namespace mesobasic
