// Device-side CSR build (graph ingest), used by pxd_initialize_edges.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Builds ranges d_row[n+1] and arcs d_nbr[2m] of the undirected edge list
// (h_fx[i], h_tx[i]), i < m (host arrays), on stream st.  Returns 0, or
// nonzero with a message in err: 1 = bad input, 2 = HIP failure.
int pxd_ingest_csr(uint32_t n, uint64_t m, const uint32_t* h_fx, const uint32_t* h_tx, uint32_t* d_row,
                   uint32_t* d_nbr, hipStream_t st, char* err, size_t errlen);
