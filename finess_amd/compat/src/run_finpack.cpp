// run_finpack.cpp -- a host frame loop for apps built against this library:
// parameters.ini -> initial data -> nout frames of FinSolveRK / FinSolveLxW ->
// q%04d.dat files.  It follows lib/2d/RunFinpack.cpp:17-130 (sampling at
// cell centres over the whole grid including ghosts, dtv carried across
// frames, tend = tstart + tfinal/nout) and the ASCII frame format of
// lib/2d/Output.cpp:26-44 (time, then q in m -> j -> i order, %24.16e).
// Frames are written by a background thread (FrameWriter), so that the text
// formatting of one frame overlaps the solve of the next.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "IniParams.h"
#include "finess_dropin.h"

#if NDIMS == 1
void QinitFunc(const dTensor1& xpts, dTensor2& qvals);
void AuxFunc(const dTensor1& xpts, dTensor2& auxvals);
#else
void QinitFunc(const dTensor2& xpts, dTensor2& qvals);
void AuxFunc(const dTensor2& xpts, dTensor2& auxvals);
#endif
#pragma weak AuxFunc

namespace {

// WriteQhelp (lib/2d/WriteQhelp.cpp:8-35, lib/1d/WriteQhelp.cpp)
void write_qhelp()
{
    const IniParams& ini = global_ini_params;
    FILE* f = std::fopen((ini.get_output_dir() + "/qhelp.dat").c_str(), "w");
    if (!f) return;
    std::fprintf(f, "%16d   : ndims   \n", ini.get_ndims());
    std::fprintf(f, "%16d   : meqn    \n", ini.get_meqn());
    std::fprintf(f, "%16d   : maux    \n", ini.get_maux());
    std::fprintf(f, "%16d   : nout    \n", ini.get_nout());
    std::fprintf(f, "%16d   : mx      \n", ini.get_mx());
#if NDIMS == 2
    std::fprintf(f, "%16d   : my      \n", ini.get_my());
#endif
    std::fprintf(f, "%16.8e : xlow    \n", ini.get_xlow());
    std::fprintf(f, "%16.8e : xhigh   \n", ini.get_xhigh());
#if NDIMS == 2
    std::fprintf(f, "%16.8e : ylow    \n", ini.get_ylow());
    std::fprintf(f, "%16.8e : yhigh   \n", ini.get_yhigh());
#endif
    std::fclose(f);
}

// Output (lib/2d/Output.cpp:13-62): q%04d.dat and a%04d.dat.  Each value
// is one "%24.16e" line (the reference's setw(24), scientific, precision 16),
// formatted with snprintf; the values (m -> j -> i order) are split into
// contiguous ranges formatted by up to 8 threads into buffers written in
// order (one 1024^2 five-field frame is 5 M lines, 1.3 s on one core).
void write_values(const std::string& path, double t, const dTensorBase& v, int nfields, int mx, int my)
{
    const long long n = (long long)nfields * mx * my;
    const int nt = (int)std::max(1LL, std::min<long long>({8LL, (long long)std::thread::hardware_concurrency(),
                                                           n / 65536 + 1}));
    std::vector<std::vector<char>> buf(nt);
    auto fmt = [&](int w) {
        const long long k0 = n * w / nt, k1 = n * (w + 1) / nt;
        std::vector<char>& b = buf[w];
        b.reserve((size_t)(k1 - k0) * 25 + 64);
        char line[64];
        auto put = [&](double x) {
            const int c = std::snprintf(line, sizeof line, "%24.16e\n", x);
            b.insert(b.end(), line, line + c);
        };
        if (w == 0) put(t);
        for (long long k = k0; k < k1; k++) {
            const int m = (int)(k / ((long long)mx * my)) + 1;
            const long long r = k % ((long long)mx * my);
            const int j = (int)(r / mx) + 1, i = (int)(r % mx) + 1;
#if NDIMS == 1
            (void)j;
            put(static_cast<const dTensorBC2&>(v).get(i, m));
#else
            put(static_cast<const dTensorBC3&>(v).get(i, j, m));
#endif
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nt; w++) th.emplace_back(fmt, w);
    fmt(0);
    for (auto& x : th) x.join();
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::cerr << "finess-hip: cannot write " << path << std::endl;
        return;
    }
    for (const auto& b : buf) std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}

void write_frame(const StateVars& Q, int n)
{
    const IniParams& ini = global_ini_params;
    std::ostringstream fa, fq;
    fa << ini.get_output_dir() << "/a" << std::setfill('0') << std::setw(4) << n << ".dat";
    fq << ini.get_output_dir() << "/q" << std::setfill('0') << std::setw(4) << n << ".dat";
#if NDIMS == 1
    const int my = 1;
#else
    const int my = ini.get_my();
#endif
    write_values(fa.str(), Q.get_t(), Q.const_ref_aux(), ini.get_maux(), ini.get_mx(), my);
    write_values(fq.str(), Q.get_t(), Q.const_ref_q(), ini.get_meqn(), ini.get_mx(), my);
}

// One frame in flight: submit() hands a copy of the state to the writer
// thread and returns at once, unless the previous frame is still being
// written, in which case it waits for that one first.
class FrameWriter {
public:
    FrameWriter() : th_([this] { loop(); }) {}
    ~FrameWriter() { finish(); }
    void submit(const StateVars& Q, int n)
    {
        std::unique_ptr<Job> j(new Job{Q, n});  // the copy, made outside the lock
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return !job_; });
        job_ = std::move(j);
        cv_.notify_all();
    }
    void finish()
    {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [this] { return !job_; });
            stop_ = true;
            cv_.notify_all();
        }
        if (th_.joinable()) th_.join();
    }

private:
    struct Job {
        StateVars Q;
        int n;
    };
    void loop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [this] { return job_ || stop_; });
            if (!job_) return;
            Job* j = job_.get();
            lk.unlock();
            write_frame(j->Q, j->n);
            lk.lock();
            job_.reset();
            cv_.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::unique_ptr<Job> job_;
    bool stop_ = false;
    std::thread th_;
};

void sample(StateVars& Q)
{
    const IniParams& ini = global_ini_params;
    const int mx = ini.get_mx(), mbc = ini.get_mbc(), meqn = ini.get_meqn(), na = std::max(1, ini.get_maux());
    const double dx = ini.get_dx(), xlow = ini.get_xlow();
#if NDIMS == 1
    const int n = mx + 2 * mbc;
    dTensor1 x(n);
    for (int i = 1 - mbc; i <= mx + mbc; i++) x.set(i + mbc, xlow + (i - 0.5) * dx);
    dTensor2 qv(n, meqn), av(n, na);
    QinitFunc(x, qv);
    if (&AuxFunc) AuxFunc(x, av); else av.setall(1.0);
    for (int i = 1 - mbc; i <= mx + mbc; i++) {
        for (int m = 1; m <= meqn; m++) Q.ref_q().set(i, m, qv.get(i + mbc, m));
        for (int m = 1; m <= na; m++) Q.ref_aux().set(i, m, av.get(i + mbc, m));
    }
#else
    const int my = ini.get_my();
    const double dy = ini.get_dy(), ylow = ini.get_ylow();
    const int nyt = my + 2 * mbc, n = (mx + 2 * mbc) * nyt;
    dTensor2 x(n, 2), qv(n, meqn), av(n, na);
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            const int k = (i + mbc - 1) * nyt + (j + mbc - 1) + 1;
            x.set(k, 1, xlow + (i - 0.5) * dx);
            x.set(k, 2, ylow + (j - 0.5) * dy);
        }
    QinitFunc(x, qv);
    if (&AuxFunc) AuxFunc(x, av); else av.setall(0.0);
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            const int k = (i + mbc - 1) * nyt + (j + mbc - 1) + 1;
            for (int m = 1; m <= meqn; m++) Q.ref_q().set(i, j, m, qv.get(k, m));
            for (int m = 1; m <= na; m++) Q.ref_aux().set(i, j, m, av.get(k, m));
        }
#endif
}

} // namespace

// main_global equivalent (lib/main_global.cpp:24-75): returns the exit code.
int finess_run(int argc, char* argv[])
{
    const std::string ini_file = argc > 1 ? argv[1] : "parameters.ini";
    try {
        global_ini_params.init(ini_file);
        const IniParams& ini = global_ini_params;
        mkdir(ini.get_output_dir().c_str(), 0755);
#if NDIMS == 1
        StateVars Q(0.0, ini.get_mx(), ini.get_meqn(), ini.get_maux(), ini.get_mbc());
#else
        StateVars Q(0.0, ini.get_mx(), ini.get_my(), ini.get_meqn(), ini.get_maux(), ini.get_mbc());
#endif
        sample(Q);
        write_qhelp();
        FrameWriter writer;
        writer.submit(Q, 0);
        finess_dropin::con_soln(Q);  // lib/2d/RunFinpack.cpp:86-89
        double dtv[3] = {0.0, ini.get_initial_dt(), ini.get_max_dt()};
        const double dtout = ini.get_tfinal() / double(ini.get_nout());
        double tend = 0.0;
        const auto t0 = std::chrono::steady_clock::now();
        for (int n = 1; n <= ini.get_nout(); n++) {
            const double tstart = tend;
            tend = tstart + dtout;
#if NDIMS == 1
            FinSolveRK(Q, tend, dtv);
#else
            if (ini.get_time_stepping_method() == IniParams::TimeSteppingMethod::LxW)
                FinSolveLxW(Q, tend, dtv);
            else if (ini.get_time_stepping_method() == IniParams::TimeSteppingMethod::USER_DEFINED)
                FinSolveUser(Q, tend, dtv);  // lib/2d/RunFinpack.cpp:117-120
            else
                FinSolveRK(Q, tend, dtv);
#endif
            writer.submit(Q, n);
            std::cout << "FINESS: Frame " << std::setw(3) << n << ": plot files done at time t ="
                      << std::setw(12) << std::scientific << std::setprecision(5) << tend << std::endl;
        }
        writer.finish();  // the last frame is on disk before the time is taken
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << " finess-hip: frames done in " << std::fixed << std::setprecision(3) << el << " s" << std::endl;
        finess_dropin::release();
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
