// CPU emulation driver (tests only): compiles the parameter kernel, the
// generic stepper and one generated register-resident stepper as plain C++
// (-DSFG_HOST_EMUL, see specfab_amd/csrc/sfg_emul.h) and runs them lane by
// lane, so the generated code's indexing and arithmetic can be checked against
// the oracle without a GPU.
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

#include "sfg_internal.h"
#include "sfg_kernels.h"

thread_local sfg_dim3 threadIdx, blockIdx, blockDim;

// A block's threads run concurrently so that workgroup barriers work; a
// thread that leaves the kernel drops out of the barrier (as a finished lane
// does on the GPU).
namespace {
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, arrived = 0;
    long gen = 0;
    void release() {
        arrived = 0;
        gen++;
        cv.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(m);
        const long g = gen;
        if (++arrived == expected) release();
        else cv.wait(lk, [&] { return gen != g; });
    }
    void drop() {
        std::lock_guard<std::mutex> lk(m);
        expected--;
        if (arrived > 0 && arrived == expected) release();
    }
};
thread_local Barrier* cur_barrier = nullptr;
}  // namespace

void sfg_emul_syncthreads() {
    if (cur_barrier) cur_barrier->wait();
}

namespace sfg {
__global__ void params_kernel(ParamArgs a);
__global__ void evolve_generic(EvolveArgs a);
void KERNEL_NAME(EvolveArgs a);
}

template <class F>
static void grid(int64_t P, int bs, F f) {
    for (int64_t b = 0; b < (P + bs - 1) / bs; b++)
        for (int t = 0; t < bs; t++) {
            blockIdx.x = (unsigned)b;
            blockDim.x = (unsigned)bs;
            threadIdx.x = (unsigned)t;
            f();
        }
}

// nblocks blocks of bs threads, each block's threads concurrent
template <class F>
static void grid_threads(int64_t nblocks, int bs, F f) {
    for (int64_t b = 0; b < nblocks; b++) {
        Barrier bar;
        bar.expected = bs;
        std::vector<std::thread> th;
        th.reserve(bs);
        for (int t = 0; t < bs; t++)
            th.emplace_back([&, t] {
                blockIdx.x = (unsigned)b;
                blockDim.x = (unsigned)bs;
                threadIdx.x = (unsigned)t;
                cur_barrier = &bar;
                f();
                cur_barrier = nullptr;
                bar.drop();
            });
        for (auto& x : th) x.join();
    }
}

// x: real SoA state (n, P), advanced in place.  which: 0 = generated kernel,
// 1 = generic kernel.  Returns 0.
extern "C" int emul_evolve(int L, int64_t P, int procs, int scheme, int nsteps, double dt, double iota, double zeta,
                           const double* ugrad, const double* gam0, const double* lam, double* x, int which,
                           int* status) {
    const sfg::RealOperator& op = sfg::real_operator(L);
    const int n = op.n, nblk = L / 2 + 1, nq = n < 15 ? n : 15;
    std::vector<double> A2(6 * 15), A4(36 * 15), tval;
    std::vector<float> tvalf;
    for (int i = 0; i < 6; i++) for (int q = 0; q < 15; q++) A2[i * 15 + q] = op.A2[i][q];
    for (int i = 0; i < 36; i++) for (int q = 0; q < 15; q++) A4[i * 15 + q] = op.A4[i][q];
    std::vector<int2> rows, grng;
    std::vector<int> row_blk, gparam, tcol;
    for (auto& r : op.rows) { rows.push_back({r.g0, r.g1}); row_blk.push_back(r.l / 2); }
    for (auto& g : op.groups) { gparam.push_back(g.param); grng.push_back({g.t0, g.t1}); }
    for (auto& t : op.terms) { tcol.push_back(t.col); tval.push_back(t.val); tvalf.push_back((float)t.val); }
    std::vector<double> theta((size_t)sfg::NPAR_K * P), omega((size_t)15 * P), sl((size_t)nblk * P);
    std::vector<double> scr((size_t)3 * n * P, 0.0);
    sfg::ParamArgs pa{};
    pa.P = P; pa.nblk = nblk; pa.nq = nq; pa.procs = procs; pa.iota = iota; pa.zeta = zeta; pa.reg_nu = op.reg_nu;
    pa.ugrad = ugrad; pa.tau = nullptr; pa.gam0 = gam0; pa.lam = lam;
    pa.A2 = A2.data(); pa.A4 = A4.data(); for (size_t b = 0; b < op.reg_shape.size() && b < 11; b++) pa.reg_shape[b] = op.reg_shape[b];
    pa.theta = theta.data(); pa.omega = omega.data(); pa.sl = sl.data();
    grid(P, 256, [&] { sfg::params_kernel(pa); });
    sfg::EvolveArgs ea{};
    ea.P = P; ea.n = n; ea.nq = nq; ea.scheme = scheme; ea.nsteps = nsteps; ea.use_ddrx = (procs & 2) ? 1 : 0;
    ea.steps = nullptr; ea.dt = dt;
    ea.rows = rows.data(); ea.row_blk = row_blk.data(); ea.grp_param = gparam.data(); ea.grp_rng = grng.data();
    ea.tcol = tcol.data(); ea.tval = tval.data(); ea.tval_f = tvalf.data();
    ea.theta = theta.data(); ea.omega = omega.data(); ea.sl = sl.data();
    ea.X = x; ea.YS = scr.data(); ea.XA = scr.data() + (size_t)n * P; ea.XB = scr.data() + (size_t)2 * n * P;
    ea.status = status; ea.zero = 0;
    if (which == 0) grid_threads((P + KERNEL_PARCELS - 1) / KERNEL_PARCELS, KERNEL_BLOCK, [&] { sfg::KERNEL_NAME(ea); });
    else grid(P, 64, [&] { sfg::evolve_generic(ea); });
    return 0;
}
