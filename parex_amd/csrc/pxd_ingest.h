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

// Work buffers of the asynchronous build: keys_in/keys_out for 2m arcs,
// deg for n + 1 counters, tmp of tmp_bytes (pxd_ingest_tmp_bytes).
struct IngestWork {
    unsigned long long *keys_in, *keys_out;
    uint32_t* deg;
    void* tmp;
    size_t tmp_bytes;
};
size_t pxd_ingest_tmp_bytes(uint32_t max_n, uint64_t max_arcs);
// The same build from device edge arrays, queued on st without a sync or an
// allocation; an end >= n sets d_bad (the arcs are then meaningless).
hipError_t pxd_ingest_csr_async(uint32_t n, uint64_t m, const uint32_t* d_fx, const uint32_t* d_tx, uint32_t* d_row,
                                uint32_t* d_nbr, const IngestWork& w, uint32_t* d_bad, hipStream_t st);
