#include "OutputQueue.hpp"

#include <memory>
#include <stdexcept>
#include <string>

#include "VtkWriter.hpp"

namespace mesobasic {

namespace {

struct Frame {
    std::vector<double> rho, u, v, fx, fy;
    std::vector<double> cap_x, cap_y;   // IB mesh (capsule models)
};

void check(int rc, const char* what)
{
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

void write_frame(const CommonParams& p, int q, int step, const Frame& f)
{
    write_model_vtk("vtkoutput", p, q, step, f.rho, f.u, f.v, f.fx, f.fy);
    if (!f.cap_x.empty()) write_capsule_vtk("vtkoutput", "capsule", step, f.cap_x, f.cap_y);
}

}  // namespace

OutputQueue::~OutputQueue()
{
    if (writer_.joinable()) writer_.join();
}

void OutputQueue::join_writer()
{
    if (writer_.joinable()) writer_.join();
    if (writer_error_) {
        std::exception_ptr e = writer_error_;
        writer_error_ = nullptr;
        std::rethrow_exception(e);
    }
}

void OutputQueue::collect_and_write(lbm_handle* h)
{
    const size_t n = static_cast<size_t>(p_.nx) * p_.ny * p_.nz;
    auto f = std::make_shared<Frame>();
    f->rho.resize(n);
    f->u.resize(n);
    f->v.resize(n);
    const bool two_d = (q_ == 9);
    if (two_d) {
        f->fx.resize(n);
        f->fy.resize(n);
    }
    int64_t taken = 0;
    check(lbm_snapshot_wait(h, &taken, f->rho.data(), f->u.data(), f->v.data(), nullptr,
                            two_d ? f->fx.data() : nullptr, two_d ? f->fy.data() : nullptr, nullptr),
          "lbm_snapshot_wait");
    pending_ = false;
    f->cap_x.swap(cap_x_);
    f->cap_y.swap(cap_y_);
    const int step = pending_step_;
    join_writer();
    if (!async_) {
        write_frame(p_, q_, step, *f);
        return;
    }
    const CommonParams* p = &p_;
    const int q = q_;
    writer_ = std::thread([this, p, q, step, f]() {
        try {
            write_frame(*p, q, step, *f);
        } catch (...) {
            writer_error_ = std::current_exception();
        }
    });
}

void OutputQueue::output(lbm_handle* h, int comp, int step)
{
    if (pending_) collect_and_write(h);
    if (capsule_) {
        // the mesh is a few values: read now, at the step of the snapshot
        const int nn = lbm_capsule_nodes(h);
        cap_x_.resize(static_cast<size_t>(nn));
        cap_y_.resize(static_cast<size_t>(nn));
        check(lbm_capsule_get_nodes(h, cap_x_.data(), cap_y_.data(), nullptr, nullptr, nullptr, nullptr),
              "lbm_capsule_get_nodes");
    }
    check(lbm_snapshot_begin(h, comp, LBM_FIELDS_REFERENCE), "lbm_snapshot_begin");
    pending_ = true;
    pending_step_ = step;
    if (!async_) collect_and_write(h);
}

void OutputQueue::finish(lbm_handle* h)
{
    if (pending_) collect_and_write(h);
    join_writer();
    // the steps after the last output are still queued on the device
    check(lbm_synchronize(h), "lbm_synchronize");
}

}  // namespace mesobasic
