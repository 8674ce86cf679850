/* Exports the reference's OpenCL inplace_fnnls (compiled as host C where it
 * lies in the reference tree, see oracle/Makefile) through a C ABI.  The same
 * file is built twice: data_type float (as shipped) and data_type double.
 * Test infrastructure only. */
#include <stddef.h>
#ifdef REF_FNNLS_F64
typedef double ref_t;
#define REF_SYM ref_inplace_fnnls_cl_f64
#else
typedef float ref_t;
#define REF_SYM ref_inplace_fnnls_cl_f32
#endif
void inplace_fnnls(ref_t const *A, ref_t const *b, ref_t *x, double const eps,
                   unsigned int const max_iterations);
/* A row-major 10x10, x in/out (the facade zeroes x before the call). */
void REF_SYM(const ref_t *A, const ref_t *b, ref_t *x, double eps, unsigned max_iter)
{
    inplace_fnnls(A, b, x, eps, max_iter);
}
