#include "BatchServer.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <dirent.h>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "../../include/mesobasic_lbm.h"
#include "InputParams.hpp"
#include "LBApp.hpp"
#include "ScmpModel.hpp"
#include "VtkWriter.hpp"

namespace mesobasic {

namespace {

using Clock = std::chrono::steady_clock;

double secs_since(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// MesoExecute.cpp:26-29 for one run directory: mkdir -p vtkoutput and
// remove the .vtk files a previous run left there.
void prepare_output_dir(const std::string& vtkdir)
{
    mkdir(vtkdir.c_str(), 0755);
    if (DIR* d = opendir(vtkdir.c_str())) {
        while (dirent* e = readdir(d)) {
            const std::string n = e->d_name;
            if (n.size() > 4 && n.compare(n.size() - 4, 4, ".vtk") == 0)
                std::remove((vtkdir + "/" + n).c_str());
        }
        closedir(d);
    }
}

struct Run {
    std::string dir;
    CommonParams p{};
    int q = 19;
    ScmpKeys k{};
    int nstep = 0, out_interval = 1;
    std::vector<double> rho0;
    int64_t id = -1;
    bool admitted = false, done = false, failed = false;
    int outputs = 0;
    std::string error;
    double t_admit = 0.0, t_done = 0.0;
};

// One thread that formats and writes the VTK files in order.
class Writer {
public:
    Writer() : th_([this] { loop(); }) {}
    ~Writer() { drain(); }

    void push(std::function<void()> task)
    {
        {
            std::lock_guard<std::mutex> g(m_);
            q_.push_back(std::move(task));
        }
        cv_.notify_one();
    }

    // Waits for every queued file; stops the thread.
    void drain()
    {
        if (!th_.joinable()) return;
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_one();
        th_.join();
    }

private:
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    bool stop_ = false;
    std::thread th_;

    void loop()
    {
        for (;;) {
            std::function<void()> task;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                task = std::move(q_.front());
                q_.pop_front();
            }
            task();
        }
    }
};

// Reads dir/input.dat as the driver would for an LBApp run of scmp3D or
// scmp2D; throws for what the batch engine does not run.
void load_run(Run& r)
{
    InputParams in(r.dir + "/input.dat");
    bool lb = false;
    for (const std::string& sec : in.section_names()) {
        if (in((sec + "app").c_str(), "false") != std::string("true")) continue;
        if (sec != "LBApp/") throw std::runtime_error("section " + sec + " asks for an app the server does not run");
        lb = true;
    }
    if (!lb) throw std::runtime_error("no [LBApp] section with app = true");
    const std::string type = in("LBApp/type", "mcmp2D");
    if (type == "scmp3D") r.q = 19;
    else if (type == "scmp2D") r.q = 9;
    else
        throw std::runtime_error("LBApp/type " + type +
                                 " is not served by the batch engine (scmp3D, scmp2D); run it with meso_amd input.dat");
    r.p = read_common_params(in);
    r.k = ScmpKeys::read(r.p, in, r.q);
    if (in("LBApp/walls", 0) != 0) throw std::runtime_error("LBApp/walls: the batch engine runs periodic lattices");
    if (in("LBApp/order", r.q == 19 ? 3 : 2) != (r.q == 19 ? 3 : 2))
        throw std::runtime_error("LBApp/order: batch jobs follow the scmp3D order (D3Q19) or the LBfluid2D order (D2Q9)");
    r.nstep = in("Time/nstep", 0);
    if (r.nstep < 0) throw std::runtime_error("Time/nstep < 0");
    const int numOutputs = in("Output/numOutputs", 1);
    r.out_interval = (numOutputs != 0) ? r.nstep / numOutputs : r.nstep + 1;
    if (r.out_interval <= 0) r.out_interval = 1;   // as the driver: numOutputs > nstep writes every step
    r.rho0 = r.k.initial_rho(r.p, r.q);
}

lbm_job_desc job_of(const Run& r, int64_t steps)
{
    lbm_job_desc d{};
    d.q = r.q;
    d.nx = r.p.nx;
    d.ny = r.p.ny;
    d.nz = (r.q == 19) ? r.p.nz : 1;
    d.tau = r.k.tauA;
    d.G = r.k.G;
    d.steps = steps;
    return d;
}

class Server {
public:
    Server(std::vector<Run>& runs, const ServeOptions& opt) : runs_(runs), opt_(opt) {}
    ~Server()
    {
        if (b_) lbm_batch_destroy(b_);
    }

    int run()
    {
        t0_ = Clock::now();
        int loaded = 0;
        for (const Run& r : runs_) loaded += r.failed ? 0 : 1;
        if (loaded == 0) {
            std::fprintf(stderr, "meso_amd --serve: no run to serve (%zu failed)\n", runs_.size());
            return static_cast<int>(runs_.size());
        }
        double shift = opt_.rho_shift;
        for (Run& r : runs_)
            if (shift < 0 && !r.failed) shift = r.k.rhoAi;
        lbm_batch_desc bd{};
        bd.precision = opt_.precision;
        bd.rho_shift = shift < 0 ? 0.5 : shift;
        bd.hbm_budget = opt_.hbm_budget;
        bd.steps_per_tick = opt_.steps_per_tick;
        bd.use_graph = 1;
        bd.autotune = 1;
        check(lbm_batch_create(&bd, &b_), "lbm_batch_create");
        if (opt_.work_limit > 0) check(lbm_batch_set_work_limit(b_, opt_.work_limit), "lbm_batch_set_work_limit");
        if (opt_.max_wait > 0) check(lbm_batch_set_max_wait(b_, opt_.max_wait), "lbm_batch_set_max_wait");

        size_t next = 0;
        int active = 0;
        int64_t lattice_updates = 0;
        for (;;) {
            // admit in input order while the budget allows
            while (next < runs_.size()) {
                Run& r = runs_[next];
                if (r.failed) {
                    next++;
                    continue;
                }
                const int rc = admit(r);
                if (rc == LBM_ENOMEM) {
                    if (active == 0) {
                        fail(r, "does not fit the HBM budget even alone");
                        next++;
                        continue;
                    }
                    break;
                }
                next++;
                if (r.done || r.failed) continue;
                active++;
            }
            if (active == 0 && next >= runs_.size()) break;
            int nfin = 0;
            check(lbm_batch_tick(b_, &nfin), "lbm_batch_tick");
            if (nfin == 0) continue;
            std::vector<int64_t> ids(static_cast<size_t>(nfin));
            int n = 0;
            check(lbm_batch_finished(b_, ids.data(), nfin, &n), "lbm_batch_finished");
            for (int i = 0; i < n; i++) {
                Run& r = *by_id(ids[i]);
                lbm_job_status st{};
                check(lbm_batch_job_status(b_, r.id, &st), "lbm_batch_job_status");
                if (st.state == LBM_JOB_FAILED) {
                    (void)lbm_batch_cancel(b_, r.id);
                    fail(r, "non-finite mass after step " + std::to_string(st.steps_done) + " (finite after step " +
                                std::to_string(st.steps_finite) + ")");
                    active--;
                    continue;
                }
                const int64_t done = st.steps_done;
                if (done % r.out_interval == 0) output(r, static_cast<int>(done));
                if (done < r.nstep) {
                    check(lbm_batch_extend(b_, r.id, std::min<int64_t>(r.out_interval, r.nstep - done)),
                          "lbm_batch_extend");
                } else {
                    check(lbm_batch_cancel(b_, r.id), "lbm_batch_cancel");
                    r.done = true;
                    r.t_done = secs_since(t0_);
                    lattice_updates += static_cast<int64_t>(r.rho0.size()) * r.nstep;
                    report(r);
                    active--;
                }
            }
        }
        writer_.drain();
        int failed = write_errors_.load();
        for (const Run& r : runs_) failed += r.failed ? 1 : 0;
        const double t = secs_since(t0_);
        std::fprintf(stderr, "meso_amd --serve: %zu runs (%d failed), %lld lattice updates in %.3f s: %.3f GLUPS incl. output\n",
                     runs_.size(), failed, static_cast<long long>(lattice_updates), t,
                     t > 0 ? lattice_updates / t / 1e9 : 0.0);
        return failed;
    }

private:
    std::vector<Run>& runs_;
    const ServeOptions& opt_;
    lbm_batch* b_ = nullptr;
    Clock::time_point t0_;
    std::atomic<int> write_errors_{0};
    Writer writer_;

    static void check(int rc, const char* what)
    {
        if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
    }

    Run* by_id(int64_t id)
    {
        for (Run& r : runs_)
            if (r.id == id) return &r;
        throw std::runtime_error("batch reported an unknown job");
    }

    void fail(Run& r, const std::string& why)
    {
        r.failed = true;
        r.error = why;
        std::printf("%s: failed: %s\n", r.dir.c_str(), why.c_str());
        std::fflush(stdout);
    }

    void report(const Run& r)
    {
        if (opt_.quiet) return;
        std::printf("%s: %s %dx%dx%d, %d steps, %d outputs, admitted %.3f s, done %.3f s\n", r.dir.c_str(),
                    r.q == 19 ? "scmp3D" : "scmp2D", r.p.nx, r.p.ny, r.q == 19 ? r.p.nz : 1, r.nstep, r.outputs,
                    r.t_admit, r.t_done);
        std::fflush(stdout);
    }

    // Submits r with steps to its first stop and writes its step-0 output.
    int admit(Run& r)
    {
        const int64_t first = std::max<int64_t>(1, std::min<int64_t>(r.out_interval, r.nstep));
        const lbm_job_desc d = job_of(r, first);
        int64_t id = -1;
        const int rc = lbm_batch_submit(b_, &d, r.rho0.data(), &id);
        if (rc == LBM_ENOMEM) return rc;
        if (rc != LBM_OK) {
            fail(r, std::string("lbm_batch_submit: ") + lbm_last_error());
            return rc;
        }
        r.id = id;
        r.admitted = true;
        r.t_admit = secs_since(t0_);
        prepare_output_dir(r.dir + "/vtkoutput");
        output(r, 0);   // MesoExecute.cpp:95-98: every app writes step 0
        if (r.nstep == 0) {
            check(lbm_batch_cancel(b_, r.id), "lbm_batch_cancel");
            r.done = true;
            r.t_done = secs_since(t0_);
            report(r);
        }
        return LBM_OK;
    }

    void output(Run& r, int step)
    {
        const size_t n = r.rho0.size();
        auto f = std::make_shared<std::vector<double>>(5 * n);
        double* a = f->data();
        const bool two_d = (r.q == 9);
        check(lbm_batch_read(b_, r.id, a, a + n, a + 2 * n, nullptr, two_d ? a + 3 * n : nullptr,
                             two_d ? a + 4 * n : nullptr, nullptr),
              "lbm_batch_read");
        r.outputs++;
        const CommonParams p = r.p;
        const int q = r.q;
        const std::string vtkdir = r.dir + "/vtkoutput";
        std::atomic<int>* errors = &write_errors_;
        writer_.push([f, p, q, n, step, vtkdir, errors]() {
            const auto& v = *f;
            std::vector<double> rho(v.begin(), v.begin() + n), u(v.begin() + n, v.begin() + 2 * n),
                vv(v.begin() + 2 * n, v.begin() + 3 * n), fx, fy;
            if (q == 9) {
                fx.assign(v.begin() + 3 * n, v.begin() + 4 * n);
                fy.assign(v.begin() + 4 * n, v.begin() + 5 * n);
            }
            try {
                write_model_vtk(vtkdir, p, q, step, rho, u, vv, fx, fy);
            } catch (const std::exception& e) {
                std::fprintf(stderr, "meso_amd --serve: %s\n", e.what());
                errors->fetch_add(1);
            }
        });
    }
};

}  // namespace

int serve_runs(const std::vector<std::string>& dirs, const ServeOptions& opt)
{
    std::vector<Run> runs(dirs.size());
    for (size_t i = 0; i < dirs.size(); i++) {
        runs[i].dir = dirs[i];
        try {
            load_run(runs[i]);
        } catch (const std::exception& e) {
            runs[i].failed = true;
            runs[i].error = e.what();
            std::printf("%s: failed: %s\n", dirs[i].c_str(), e.what());
        }
    }
    Server s(runs, opt);
    return s.run();
}

}  // namespace mesobasic
