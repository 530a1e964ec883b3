// Build-time helper: writes the sparsity structure of the real-basis operator
// of one truncation L (rows -> parameter groups -> (column, term index)) as
// text, for specfab_amd/tables/gen_fast.py, which compiles the term values
// into the generated kernels as literal constants.
//
//   dump_basis L > basis_L.txt
// Format:  "L n nrows ngroups nterms"  then per row "R l ngroups", per group
// "G param nterms", per term "T col termindex value" (value as a C hex float).
#include <cstdio>
#include <cstdlib>

#include "../csrc/sfg_internal.h"

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    int L = std::atoi(argv[1]);
    const sfg::RealOperator& op = sfg::real_operator(L);
    std::printf("%d %d %zu %zu %zu\n", op.L, op.n, op.rows.size(), op.groups.size(), op.terms.size());
    for (const auto& r : op.rows) {
        std::printf("R %d %d\n", r.l, r.g1 - r.g0);
        for (int g = r.g0; g < r.g1; g++) {
            const auto& G = op.groups[g];
            std::printf("G %d %d\n", G.param, G.t1 - G.t0);
            for (int t = G.t0; t < G.t1; t++) std::printf("T %d %d %a\n", op.terms[t].col, t, op.terms[t].val);
        }
    }
    return 0;
}
