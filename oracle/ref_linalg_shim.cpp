// Exports the reference's Eigen-free linalgebra templates (compiled where they
// lie in the reference tree, see oracle/Makefile) through a C ABI so that the
// golden-vector script can record their outputs.  Test infrastructure only.
#include "linalgebra/interface/cholesky.hpp"
#include "linalgebra/interface/common.hpp"
#include "linalgebra/interface/fused_cholesky_substitution.hpp"
#include "linalgebra/interface/substitution.hpp"

#define EXPORT_ALL(T, sfx)                                                        \
    extern "C" void ref_cholesky_decomp_##sfx(T *M, T *L, int full, int view) {   \
        cholesky_decomp<T>(M, L, full, view);                                     \
    }                                                                             \
    extern "C" void ref_forward_##sfx(T *L, T *b, T *y, int full, int view) {     \
        solve_forward_substitution<T>(L, b, y, full, view);                       \
    }                                                                             \
    extern "C" void ref_backward_##sfx(T *L, T *y, T *x, int full, int view) {    \
        solve_backward_substitution<T>(L, y, x, full, view);                      \
    }                                                                             \
    extern "C" void ref_swap_row_column_##sfx(T *M, int i, int j, int full,       \
                                              int view) {                         \
        swap_row_column<T>(M, i, j, full, view);                                  \
    }                                                                             \
    extern "C" void ref_rcaddition_##sfx(T *M, T *L, T *b, T *y, int full,        \
                                         int view) {                              \
        fused_cholesky_forward_substitution_solver_rcaddition<T>(M, L, b, y,      \
                                                                 full, view);     \
    }                                                                             \
    extern "C" void ref_inbetween_removal_##sfx(T *M, T *L, T *b, T *y, int pos,  \
                                                int full, int view) {             \
        fused_cholesky_forward_substitution_solver_inbetween_removal<T>(          \
            M, L, b, y, pos, full, view);                                         \
    }

EXPORT_ALL(double, f64)
EXPORT_ALL(float, f32)
