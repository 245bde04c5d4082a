// Device tables of the folded potential used by evaluation.
#pragma once

#include "device_tables.hpp"

namespace plm {

struct EvalTypeDev {
    int n_heads;            // non-conjugate components (gtinv) or pair components
    const int* pot_off;     // [n_heads+1] first term of each head
    // The terms (coefficients and the order-parameter and feature keys of their
    // products, ordered by head) for the per-atom contraction, dealt to the
    // pq_lanes lanes of a workgroup in equal contiguous chunks of pq_iters terms
    // and interleaved: record i * pq_lanes + l is the i-th term of lane l's
    // chunk (past the chunk: padding, head 0xffff). Keys and head in 16-bit fields,
    // 0xffff: none:
    //   (pd0 | pd1 << 16, pd2 | head << 16, pf0 | pf1 << 16, 0), coefficients (e, f).
    int pq_iters;
    const int4* pq_key;
    const double2* pq_c;
    // half-pair contraction: heads of this type grouped by type pair
    const int* hd_off;      // [n_tp+1]
    // gtinv: the heads of each type pair are ordered by Y key; hd_koff[tp (n_lmh + 1) + k]
    // is the first head of key k of type pair tp
    const int* hd_koff;
    // per head, one 16-byte record: (radial n_id, ylm key, ilocal_noconj of the
    // head (atom i), jlocal_noconj (the same component seen from atom j) with
    // bit 28 set where (-1)^l = -1 and bit 29 where m == 0 (weight 1/2))
    const int4* hd_rec;
};

struct EvalDev {
    EvalTypeDev t[kMaxTypes];
    int pq_lanes;           // lanes of the contraction's workgroup: 64 or 256
};

}  // namespace plm
