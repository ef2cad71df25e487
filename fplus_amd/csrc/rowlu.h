// rowlu.h -- the row kernel's host side (rowlu.cpp): the device elimination of a CSR matrix, and the compressed
// transposes its pipeline uses.
#pragma once
#include "fplus.h"

#include <cstdint>
#include <vector>

namespace fpa {

// The row kernel on the device (reference src/factor_row.c:1191-1685): the CSR matrix (pattern Rp / Ri on the host,
// canonical; values d_Rx in device memory, vs doubles each) factored with full (0), partial (1) or rook (2) pivoting
// and, for drop_tol > 0, the incomplete factorization's drops.  out receives CSR factors (csc = 0): L with its unit
// diagonal first in each row, U from the diagonal on, p / q the input row / column at each position.
int row_factor_device(int64_t n, const int64_t* Rp, const int64_t* Ri, const double* d_Rx, int vs, double drop_tol, int strategy,
                      double pivot_rel, int verbose, fplus_factors* out, char* err, size_t errlen);

// the transpose of a compressed n x n matrix (CSC <-> CSR of the same matrix), indices ascending in every line; map[t]
// = the input position of output entry t
void compressed_transpose(int64_t n, const int64_t* P, const int64_t* I, std::vector<int64_t>& TP, std::vector<int64_t>& TI,
                          std::vector<int64_t>& map);

// row-compressed factors to column-compressed ones, in place (reference src/pipeline.c:152-166)
void factors_to_csc(fplus_factors* f);

}  // namespace fpa
