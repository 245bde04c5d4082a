#pragma once

#include <vector>

namespace plm {

// Lattice used for the translations (the reduced cell when reduction
// happened), the index box of the translation search and, for each integer
// shift (i, j, k) in that box, its index in the translation list or -1.
struct CellInfo {
    double axis[9];
    int mexp[3];
    std::vector<int> tindex;
};

// axis9: row-major 3x3 with lattice vectors as columns; pos_c: 3 x n
// (component-major). Outputs translations (n_T x 3) and the positions used
// for the neighbour search (rewrapped only when the cell was reduced), 3 x n.
void cell_translations(const double* axis9, const double* pos_c, int n, double cutoff,
                       std::vector<double>& trans, std::vector<double>& pos_out,
                       CellInfo* info = nullptr);

}  // namespace plm
