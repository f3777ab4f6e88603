#include "VtkWriter.hpp"

#include <algorithm>
#include <charconv>
#include <exception>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace mesobasic {

// v as printf("%.3f") prints it in the C locale (std::to_chars in fixed form
// with a precision is specified to give exactly that, "nan", "-nan", "inf"
// and "-0.000" included).  Needs kMaxFixed3 bytes at p.
char* put_fixed3(char* p, double v)
{
    return std::to_chars(p, p + kMaxFixed3, v, std::chars_format::fixed, 3).ptr;
}

namespace {

// Formats lines [0, nlines) into text, line l by emit(l, p) -> end pointer,
// each line at most max_line bytes; large files are formatted by several
// threads (each its own block of lines, joined in order).  The reference
// ends every line with std::endl, which flushes the stream per value; the
// bytes are the same.
template <typename F>
std::string format_lines(long long nlines, size_t max_line, F emit)
{
    const long long kMinPerThread = 1 << 16;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = static_cast<int>(std::max(1LL, std::min<long long>(hw, nlines / kMinPerThread)));
    std::vector<std::string> parts(nt);
    auto work = [&](int t) {
        const long long l0 = nlines * t / nt, l1 = nlines * (t + 1) / nt;
        std::string& s = parts[t];
        s.resize(static_cast<size_t>(l1 - l0) * 8 + max_line);
        size_t at = 0;
        for (long long l = l0; l < l1; l++) {
            if (s.size() - at < max_line) s.resize(std::max(s.size() * 2, at + max_line));
            at = static_cast<size_t>(emit(l, &s[at]) - s.data());
        }
        s.resize(at);
    };
    if (nt == 1) {
        work(0);
        return std::move(parts[0]);
    }
    // an exception in a worker (an allocation failing) is rethrown here
    std::vector<std::exception_ptr> err(nt);
    auto guarded = [&](int t) {
        try {
            work(t);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(guarded, t);
    guarded(0);
    for (auto& t : th) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    size_t total = 0;
    for (auto& s : parts) total += s.size();
    std::string out;
    out.reserve(total);
    for (auto& s : parts) out += s;
    return out;
}

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
    // loop order and selection of LBfluid3D.cpp:399-414 (one rank, xOffset
    // 0): k, then j, then the i that iskip divides, x fastest
    const int nk = (NZ + kskip - 1) / kskip, nj = (NY + jskip - 1) / jskip, ni = NX / iskip;
    const std::string text = format_lines(
        static_cast<long long>(nk) * nj * ni, kMaxFixed3 + 1, [&](long long l, char* p) {
            const int i = static_cast<int>(l % ni + 1) * iskip;
            const long long kj = l / ni;
            const int j = static_cast<int>(kj % nj) * jskip + 1, k = static_cast<int>(kj / nj) * kskip + 1;
            const long long c = (static_cast<long long>(i - 1) * NY + (j - 1)) * NZ + (k - 1);
            p = put_fixed3(p, rho[c]);
            *p++ = '\n';
            return p;
        });
    out.write(text.data(), static_cast<std::streamsize>(text.size()));
    out.flush();
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
    // LBfluid2D.cpp:440-453: j, then the i that iskip divides, x fastest
    const int nj = (NY + jskip - 1) / jskip, ni = NX / iskip;
    const long long nl = static_cast<long long>(nj) * ni;
    auto node = [&](long long l) {
        const int i = static_cast<int>(l % ni + 1) * iskip, j = static_cast<int>(l / ni) * jskip + 1;
        return static_cast<long long>(i - 1) * NY + (j - 1);
    };
    std::string text = format_lines(nl, kMaxFixed3 + 1, [&](long long l, char* p) {
        p = put_fixed3(p, rho[node(l)]);
        *p++ = '\n';
        return p;
    });
    out.write(text.data(), static_cast<std::streamsize>(text.size()));
    // LBfluid2D.cpp:461-479
    out << "   " << std::endl;
    out << "VECTORS Velocity float" << std::endl;
    text = format_lines(nl, 3 * kMaxFixed3 + 3, [&](long long l, char* p) {
        const long long c = node(l);
        p = put_fixed3(p, ustar[c]);
        *p++ = ' ';
        p = put_fixed3(p, vstar[c]);
        *p++ = ' ';
        p = put_fixed3(p, 0.0);
        *p++ = '\n';
        return p;
    });
    out.write(text.data(), static_cast<std::streamsize>(text.size()));
    out.flush();
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
