// Homogenization kernel arguments and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfg {

struct EijArgs {
    int64_t P;
    int nq;
    int frame;            // 0 = a2 eigenframe, 1 = given frames, 2 = Cartesian (x, y, z)
    int n_grain;          // 1 or -3 (3: Taylor part only, the Sachs part by eij_sachs3_kernel)
    double Ecc, Eca, alpha;
    const double* x;      // real SoA state (n, P)
    const double* A2;     // (6, 15)
    const double* A4;     // (36, 15)
    const double* frames; // (P, 3, 3) rows e1, e2, e3 (frame == 1)
    double* Eij;          // (P, 6)
    double* lami;         // (P, 3) or null
    double* ei;           // (P, 3, 3) or null
    int* status;          // (P) or null
    // optional work list: items t < nidx evaluate state column idx[t] and
    // write Eij/lami/ei at row t (status stays at the parcel's own row);
    // with nidx_dev the item count is min(*nidx_dev, nidx), read on the
    // device (nidx is then the list's capacity), so no host round trip
    const int* idx;
    int64_t nidx;
    const int* nidx_dev;
    // batch form (idx null): one flag per wavefront of 64 items, written by
    // the first pass (1: one of its rows needs the regularized second pass),
    // or null (the second pass then scans the rows)
    int* wave_flag;
    // Evw (launch_evw): nstates stress states (v, w, tau) given directly,
    // states (nstates, 15) = v(3), w(3), tau(9) row-major (tau symmetric);
    // Eij then holds (P, nstates)
    const double* states;
    int nstates;
    // filled by launch_eij from (Ecc, Eca, n_grain): rheo_params_tranisotropic
    // (src/rheologies.f90:123-132) of the Sachs (ef = +1) and Taylor (ef = -1)
    // branches and the isotropic normalisations (src/homogenizations.f90:207, 235)
    double sA, sB, sC, tA, tB, tC, sachs_iso, taylor_iso;
};

hipError_t launch_eij(const EijArgs& a, hipStream_t s);

// n' = 3 Sachs (src/homogenizations.f90:93-103): the full-a4 map (15 x 15,
// the reference body) and its isotropic entries, the state's row count
struct Sachs3Args {
    const double* A4F;
    double u_iso[15];
    int nrows;
};
// Eij with n' = 3 grains: a.alpha > 0 mixes with the Taylor values that
// launch_eij (n_grain = 3, alpha = 1) left in a.Eij; the Sachs parameters are
// filled here from (Ecc, Eca)
hipError_t launch_eij_sachs3(const EijArgs& a, const Sachs3Args& b, hipStream_t s);
// Evw of the given stress states (a.states, a.nstates) per parcel
// (src/enhancementfactors.f90:47-69): n_grain 1 or -3 (Sachs + Taylor), or
// the Taylor part with n' = 3 parameters (n_grain 3, alpha = 1) that
// launch_evw_sachs3 then mixes with the n' = 3 Sachs part
hipError_t launch_evw(const EijArgs& a, hipStream_t s);
hipError_t launch_evw_sachs3(const EijArgs& a, const Sachs3Args& b, hipStream_t s);
// <D> with n_RSS = 4 (src/dynamics.f90:424-449): x (nrows, P), tau (P, 3, 3) -> out (P)
hipError_t launch_ev_D4(int64_t P, int nrows, const double* A4F, const double* x, const double* tau, double* out,
                        hipStream_t s);
// E_CAFFE (src/enhancementfactors.f90:302-332) of <D> (P), in place
// exact sixth / eighth moments (P, 28) / (P, 45), monomials <x^a y^b z^c> in
// the order a = d..0, b = d - a..0, of states with nrows rows (rows above
// l = 8 do not enter)
hipError_t launch_moments68(int64_t P, int nrows, const double* x, double* m6, double* m8, hipStream_t s);
// eigenpairs of P symmetric 3x3 matrices (P, 9): ei (P, 9) rows, lami (P, 3), largest first
hipError_t launch_eig3(int64_t P, const double* M, double* ei, double* lami, hipStream_t s);
hipError_t launch_caffe(int64_t P, int n_rss, double Emin, double Emax, double* DE, hipStream_t s);

}  // namespace sfg
