// seams_internal.h -- pieces of seams.hip shared with other translation
// units of the library (internal C++ API, not part of the C ABI).
#pragma once

#include <string>

#include "kernels.h"

namespace smi {
namespace seams {

// Records the message for smi_last_error() and returns code.
int op_fail(int code, const std::string &m);

// Geometry of the per-device entry points, with the checks of smi_create.
int make_geom(int ndim, int order, const int n_cells[3], const int patch_size[3], const double *cell_length,
              Geom &g);

// Scratch of one sort/exchange call, carved out of one buffer of 4-byte
// words (each part aligned to 256 bytes).
struct SortScratch {
    SortBufs sb;
    uint32_t *keys;
    double *stage;
    int *err;
};

struct SortArrays {
    double *x, *y, *z, *px, *py, *pz, *w;
    int16_t *charge;
    int32_t *key, *patch_count, *bin_index;
};

// 4-byte words of scratch a sort/exchange of n particles over geometry g
// needs, and its carving out of a buffer of that many words.
int64_t sort_scratch_words(const Geom &g, int64_t n);
SortScratch carve_sort_scratch(const Geom &g, int64_t n, int32_t *base);

// Checks of smi_sort_exchange's arguments; fills g.
int sort_exchange_args(int ndim, const int n_cells[3], const int patch_size[3], const double cell_length[3],
                       int64_t n, const SortArrays &A, Geom &g);

// The sort/exchange on scratch s, all on stream st (counts, placed-stayer
// flags and the error flag are cleared first, so the scratch may hold
// anything on entry); returns after the stream has drained.
int sort_exchange_impl(const Geom &g, int64_t n, const SortArrays &A, SortScratch &s, hipStream_t st);

}  // namespace seams
}  // namespace smi
