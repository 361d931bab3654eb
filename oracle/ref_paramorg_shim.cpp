// ref_paramorg_shim.cpp -- C entry points around the reference's own
// ParameterOrganizer<double> (interface/parameterorganizer/parameterorganizer.cpp),
// compiled from the reference sources by oracle/build_ref.sh into
// oracle/_ref/libref_paramorg.so.  TEST INFRASTRUCTURE ONLY: it pins the
// oracle's knot packing (orc_organize_in / orc_organize_out) against the
// reference code itself.
#include "interface/parameterorganizer/parameterorganizer.h"

extern "C" {

// organize2D1(px_in, py_in, p_out, ...) (parameterorganizer.cpp:30-55).
void ref_organize_in(const double *px, const double *py, int n_frames, int n_blocks, double *p) {
    int nk = n_blocks + 1, npp = n_frames * n_blocks + 1;
    NeutralMatrix<double> mx(const_cast<double *>(px), nk, n_frames, false);
    NeutralMatrix<double> my(const_cast<double *>(py), nk, n_frames, false);
    NeutralMatrix<double, int, RowMajor<int>> out;
    ParameterOrganizer<double>::organize2D1(mx, my, out, n_frames, n_blocks, npp);
    for (int i = 0; i < 2 * npp; i++) p[i] = out[i];
}

// organize1D2(p_in, px_out, py_out, ...) (parameterorganizer.cpp:79-99).
void ref_organize_out(const double *p, int n_frames, int n_blocks, double *px, double *py) {
    int nk = n_blocks + 1, npp = n_frames * n_blocks + 1;
    NeutralMatrix<double> mp(const_cast<double *>(p), 2 * npp, 1, false);
    NeutralMatrix<double> ox(px, nk * n_frames, 1, false);
    NeutralMatrix<double> oy(py, nk * n_frames, 1, false);
    ParameterOrganizer<double>::organize1D2(mp, ox, oy, n_frames, n_blocks, npp);
}
}
