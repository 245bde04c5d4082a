// Device tables of the folded potential used by evaluation.
#pragma once

#include "device_tables.hpp"

namespace plm {

struct EvalTypeDev {
    int n_heads;            // non-conjugate components (gtinv) or pair components
    const int* pot_off;     // [n_heads+1]
    const double* pot_ce;   // energy coefficient of each term
    const double* pot_cf;   // force coefficient of each term
    const int* pot_pd;      // [nterm*3] order-parameter keys of the product, -1 padded
    const int* pot_pf;      // [nterm*2] feature keys of the product of features, -1 padded
    // half-pair contraction: heads of this type grouped by type pair
    const int* hd_off;      // [n_tp+1]
    const int* hd_nid;
    const int* hd_ylm;
    const int* hd_i;        // ilocal_noconj of the head (atom i)
    const int* hd_j;        // jlocal_noconj (same component seen from atom j)
    const float* hd_signj;  // (-1)^l
    const unsigned char* hd_m0;  // m == 0 (weight 1/2)
};

struct EvalDev {
    EvalTypeDev t[kMaxTypes];
};

}  // namespace plm
