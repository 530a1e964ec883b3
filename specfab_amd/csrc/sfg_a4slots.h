// Index-sorted slots of the full a4 (15 distinct entries a0 <= a1 <= a2 <= a3,
// in the order of the reference's ev_c4 body) for every (i, j, k, l), with
// the body's a4(3,2,1,2) entry, which it sets to its a4(1,2,3,3) entry
// (src/include/ev_c4__body.f90), and the plain symmetric slots.  Tables, so
// that every use inside unrolled loops folds to a constant.
#pragma once

namespace sfg {

__host__ __device__ constexpr int a4_sorted_slot(int i, int j, int k, int l) {
    int v[4] = {i, j, k, l};
    for (int p = 0; p < 4; p++)
        for (int q = p + 1; q < 4; q++)
            if (v[q] < v[p]) {
                int t = v[p];
                v[p] = v[q];
                v[q] = t;
            }
    int idx = 0;
    for (int a0 = 0; a0 < 3; a0++)
        for (int a1 = a0; a1 < 3; a1++)
            for (int a2 = a1; a2 < 3; a2++)
                for (int a3 = a2; a3 < 3; a3++) {
                    if (a0 == v[0] && a1 == v[1] && a2 == v[2] && a3 == v[3]) return idx;
                    idx++;
                }
    return -1;
}

__host__ __device__ constexpr int a4_slot(int i, int j, int k, int l) {
    return (i == 2 && j == 1 && k == 0 && l == 1) ? 8 : a4_sorted_slot(i, j, k, l);
}

struct A4Slots {
    int s[81];     // the reference body's slots (its (3,2,1,2) entry included)
    int sym[81];   // the symmetric tensor's slots
};
__host__ __device__ constexpr A4Slots make_a4_slots() {
    A4Slots t{};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) {
                    t.s[27 * i + 9 * j + 3 * k + l] = a4_slot(i, j, k, l);
                    t.sym[27 * i + 9 * j + 3 * k + l] = a4_sorted_slot(i, j, k, l);
                }
    return t;
}
constexpr A4Slots kA4Slots = make_a4_slots();

}  // namespace sfg
