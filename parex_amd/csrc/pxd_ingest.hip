// Device-side CSR build from an undirected edge list (graph ingest).
//
// Same result as the reference's host build: every edge (u, v) gives the arcs
// u -> v and v -> u, arcs are ordered by (source, target), ranges[u] is the
// first arc of u (lib/utils/graph_io.h:49-84 reads the edges,
// lib/core/graph.h:147-184 sorts them and fills the ranges).  Loops and
// duplicate edges are kept as they come, as the reference keeps them.
//
// One 64-bit key (source << 32 | target) per arc, sorted with rocPRIM's device
// radix sort over the significant bits only; degrees by atomics and an
// exclusive scan give the ranges.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "pxd_ingest.h"

namespace pxd {

__global__ void k_edge_keys(uint64_t m, const uint32_t* __restrict__ fx, const uint32_t* __restrict__ tx,
                            uint32_t n, unsigned long long* __restrict__ keys, uint32_t* __restrict__ deg,
                            uint32_t* __restrict__ bad) {
    uint32_t b = 0u;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t u = fx[e], v = tx[e];
        if (u >= n || v >= n) {       // reported; the key stays in range
            b = 1u;
            u = v = 0u;
        }
        keys[2 * e] = ((unsigned long long)u << 32) | v;
        keys[2 * e + 1] = ((unsigned long long)v << 32) | u;
        atomicAdd(&deg[u], 1u);
        atomicAdd(&deg[v], 1u);
    }
    if (b) atomicOr(bad, 1u);
}

__global__ void k_key_targets(uint64_t arcs, const unsigned long long* __restrict__ keys, uint32_t* __restrict__ nbr) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < arcs; i += (uint64_t)gridDim.x * blockDim.x)
        nbr[i] = (uint32_t)keys[i];
}

static uint32_t grid_of_items(uint64_t items) {
    const uint64_t g = (items + 255) / 256;
    return (uint32_t)(g < 1 ? 1 : g > 8192 ? 8192 : g);
}

}  // namespace pxd

// Sort bits for sources below n: the 32 target bits and ceil(log2 n) source bits.
static int key_end_bit(uint32_t n) {
    int endbit = 32;
    while (endbit < 64 && (1ull << (endbit - 32)) < (unsigned long long)n) ++endbit;
    return endbit;
}

size_t pxd_ingest_tmp_bytes(uint32_t max_n, uint64_t max_arcs) {
    size_t tb_scan = 0, tb_sort = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)max_n + 1) !=
            hipSuccess ||
        hipcub::DeviceRadixSort::SortKeys(nullptr, tb_sort, (unsigned long long*)nullptr,
                                          (unsigned long long*)nullptr, (int)max_arcs, 0, 64) != hipSuccess)
        return 0;
    return tb_scan > tb_sort ? tb_scan : tb_sort;
}

hipError_t pxd_ingest_csr_async(uint32_t n, uint64_t m, const uint32_t* d_fx, const uint32_t* d_tx, uint32_t* d_row,
                                uint32_t* d_nbr, const IngestWork& w, uint32_t* d_bad, hipStream_t st) {
    using namespace pxd;
    const uint64_t arcs = 2 * m;
    hipError_t e;
    if ((e = hipMemsetAsync(w.deg, 0, ((size_t)n + 1) * 4, st)) != hipSuccess) return e;
    if (m) {
        k_edge_keys<<<grid_of_items(m), 256, 0, st>>>(m, d_fx, d_tx, n, w.keys_in, w.deg, d_bad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    size_t tb = w.tmp_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.deg, d_row, (int)n + 1, st)) != hipSuccess) return e;
    if (arcs) {
        tb = w.tmp_bytes;
        if ((e = hipcub::DeviceRadixSort::SortKeys(w.tmp, tb, w.keys_in, w.keys_out, (int)arcs, 0, key_end_bit(n),
                                                   st)) != hipSuccess)
            return e;
        k_key_targets<<<grid_of_items(arcs), 256, 0, st>>>(arcs, w.keys_out, d_nbr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t get(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
};
}  // namespace

int pxd_ingest_csr(uint32_t n, uint64_t m, const uint32_t* h_fx, const uint32_t* h_tx, uint32_t* d_row,
                   uint32_t* d_nbr, hipStream_t st, char* err, size_t errlen) {
    const uint64_t arcs = 2 * m;
    if (arcs > 0x7FFFFFFFull) {
        snprintf(err, errlen, "device CSR build takes below 2^31 arcs (%llu given)", (unsigned long long)arcs);
        return 1;
    }
    DevBuf fx, tx, kin, kout, deg, bad, tmp;
    hipError_t e;
#define PXI(x)                                                                     \
    do {                                                                           \
        if ((e = (x)) != hipSuccess) {                                             \
            snprintf(err, errlen, "%s: %s", #x, hipGetErrorString(e));             \
            return 2;                                                              \
        }                                                                          \
    } while (0)
    const size_t tb = pxd_ingest_tmp_bytes(n, arcs);
    if (!tb) {
        snprintf(err, errlen, "sizing the device sort failed");
        return 2;
    }
    PXI(fx.get(m * 4));
    PXI(tx.get(m * 4));
    PXI(kin.get(arcs * 8));
    PXI(kout.get(arcs * 8));
    PXI(deg.get(((size_t)n + 1) * 4));
    PXI(bad.get(4));
    PXI(tmp.get(tb));
    if (m) {
        PXI(hipMemcpyAsync(fx.p, h_fx, m * 4, hipMemcpyHostToDevice, st));
        PXI(hipMemcpyAsync(tx.p, h_tx, m * 4, hipMemcpyHostToDevice, st));
    }
    PXI(hipMemsetAsync(bad.p, 0, 4, st));
    const IngestWork w{(unsigned long long*)kin.p, (unsigned long long*)kout.p, (uint32_t*)deg.p, tmp.p, tb};
    PXI(pxd_ingest_csr_async(n, m, (const uint32_t*)fx.p, (const uint32_t*)tx.p, d_row, d_nbr, w, (uint32_t*)bad.p,
                             st));
    uint32_t hbad = 0;
    PXI(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, st));
    PXI(hipStreamSynchronize(st));    // the temporaries are freed on return
#undef PXI
    if (hbad) {
        for (uint64_t i = 0; i < m; ++i)
            if (h_fx[i] >= n || h_tx[i] >= n) {
                snprintf(err, errlen, "edge %llu (%u, %u) has an end >= n = %u", (unsigned long long)i, h_fx[i],
                         h_tx[i], n);
                break;
            }
        return 1;
    }
    return 0;
}
