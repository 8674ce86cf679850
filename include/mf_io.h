/*
 * mf_io.h -- ROOT-free reader of the reference's sample files (host code in
 * libmf_amd.so).
 *
 * The reference's drivers read their inputs with ROOT:
 *   - the `Samples` tree written by gen/src/gen_data.cpp:224-248 (samples,
 *     amplitudeTruth, ... per event), read by multifit_cpu/src/multifit.cpp:
 *     88-94 and io/src/ecal_data_io.cpp:6-12;
 *   - the pulse shape file (data/EmptyFile.root): the TGraph
 *     PulseShape/grPulseShape and the tree PulseShape/Tail (timeMin,
 *     expAmplitude, expTime), read by multifit_cpu/interface/Pulse.h:69-89;
 *   - the `RecoAndSim` tree the CPU driver writes (multifit.cpp:108-129:
 *     chi2, samplesReco, ipulseintime, ...).
 * mf_io_write_tree writes a tree of such branches to a new file.
 * This reader covers what those files hold: the file's key list, zlib
 * compressed records, tree branches of one basic value per entry (/F, /D,
 * /I) or of std::vector<double> / std::vector<int>, and TGraph points.  It
 * does not interpret streamer info: the caller names the branch's kind, as
 * the reference's SetBranchAddress calls do.
 *
 * Limits: baskets must be stored as records of their own (ROOT writes them
 * so when a tree is flushed or its file closed); LZMA, LZ4 and ZSTD
 * compressed records are reported as MF_ERR_UNSUPPORTED.  A file whose
 * header was never updated (a writer that did not close it) is read by
 * scanning its records to the end, as ROOT's own recovery does.
 */
#ifndef MF_IO_H
#define MF_IO_H

#include <stdint.h>

#include "mf_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mf_rootfile mf_rootfile;

/* value kind of a branch, as its leaf type or C++ type in the writer */
enum {
    MF_IO_F32 = 0,      /* "x/F"                  */
    MF_IO_F64 = 1,      /* "x/D"                  */
    MF_IO_I32 = 2,      /* "x/I"                  */
    MF_IO_VEC_F64 = 3,  /* std::vector<double>    */
    MF_IO_VEC_I32 = 4   /* std::vector<int>       */
};

/* Opens (memory-maps) a ROOT file and lists its records. */
mf_status mf_io_open(const char *path, mf_rootfile **file);
mf_status mf_io_close(mf_rootfile *file);

/* Entries of tree `tree` (a name, or "dir/name") as its header records
 * them; -1 when the file holds no header for it (never closed). */
mf_status mf_io_tree_entries(mf_rootfile *file, const char *tree, int64_t *entries);

/* Reads branch `branch` of tree `tree` in entry order, every value as a
 * double.  out: values, all entries concatenated (cap doubles; NULL to only
 * count); counts: values per entry (counts_cap entries; NULL to skip).
 * *nvalues / *nentries: the totals, also when they exceed the caps (the
 * call then fails with MF_ERR_INVALID after filling what fits). */
mf_status mf_io_read_branch(mf_rootfile *file, const char *tree, const char *branch,
                            int32_t kind, double *out, int64_t cap, int64_t *nvalues,
                            int64_t *counts, int64_t counts_cap, int64_t *nentries);

/* Points of TGraph `name` ("dir/name" or a name): x[i], y[i] for i < cap;
 * *n: the number of points. */
mf_status mf_io_read_graph(mf_rootfile *file, const char *name, double *x, double *y,
                           int64_t cap, int64_t *n);

/* One branch of a tree to write: its name, kind (MF_IO_*), the values of
 * all entries concatenated (as doubles; /F and /I branches store them
 * rounded to float and truncated to int), and for the vector kinds the
 * number of values of each entry. */
typedef struct mf_io_column {
    const char    *name;
    int32_t        kind;
    int32_t        pad;
    const double  *values;
    const int64_t *counts;
} mf_io_column;

/* Writes a new ROOT file at `path` holding one tree, `tree`, of nentries
 * entries and ncols branches: what the reference's CPU driver writes with
 * TTree::Branch / Fill / Write (multifit_cpu/src/multifit.cpp:107-129,
 * 266-279; its RecoAndSim tree).  The file is uncompressed, in ROOT 6's
 * small-file layout and the class versions ROOT 6 holds for TTree and its
 * branches and leaves (no streamer info needed); baskets hold up to 32000
 * bytes of entries.  Files beyond 2 GB are refused (MF_ERR_UNSUPPORTED).
 * Branches of std::vector<std::vector<double>> are not written. */
mf_status mf_io_write_tree(const char *path, const char *tree, const char *title,
                           const mf_io_column *cols, int32_t ncols, int64_t nentries);

#ifdef __cplusplus
}
#endif
#endif
