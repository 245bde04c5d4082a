#pragma once

#include <vector>

namespace plm {

// axis9: row-major 3x3 with lattice vectors as columns; pos_c: 3 x n
// (component-major). Outputs translations (n_T x 3) and the positions used
// for the neighbour search (rewrapped only when the cell was reduced), 3 x n.
void cell_translations(const double* axis9, const double* pos_c, int n, double cutoff,
                       std::vector<double>& trans, std::vector<double>& pos_out);

}  // namespace plm
