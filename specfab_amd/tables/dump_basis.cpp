// Build-time helper: writes the sparsity structure of the real-basis operator
// of one truncation L (rows -> parameter groups -> (column, term index)) as
// text, for specfab_amd/tables/gen_fast.py, which compiles the term values
// into the generated kernels as literal constants.
//
//   dump_basis L > basis_L.txt
// Format:  "L n nrows ngroups nterms"  then per row "R l ngroups", per group
// "G param nterms", per term "T col termindex value" (value as a C hex float).
//
//   dump_basis maps
// writes the nonzero coefficients of the Mandel a2 and a4 maps of the l <= 4
// state block (a2v = A2 x / x0 + I/3, a4v = A4 x / x0) as "A2 i q value" and
// "A4 i q value" lines (i: Mandel entry, 6 resp. 36; q: real-basis row).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../csrc/sfg_internal.h"

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    if (std::strcmp(argv[1], "maps") == 0) {
        const sfg::RealOperator& op4 = sfg::real_operator(4);
        for (int i = 0; i < 6; i++)
            for (int q = 0; q < 15; q++)
                if (op4.A2[i][q] != 0.0) std::printf("A2 %d %d %a\n", i, q, op4.A2[i][q]);
        for (int i = 0; i < 36; i++)
            for (int q = 0; q < 15; q++)
                if (op4.A4[i][q] != 0.0) std::printf("A4 %d %d %a\n", i, q, op4.A4[i][q]);
        return 0;
    }
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
